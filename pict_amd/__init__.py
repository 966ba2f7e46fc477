"""pict_amd: the PISO time step of PICT on MI355X (gfx950) behind the reference's
PISOtorch / PISOtorch_diff / PISOtorch_simulation Python API."""
