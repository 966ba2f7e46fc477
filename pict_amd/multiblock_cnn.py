"""Convolutions over multi-block domains: `MultiblockConv`, the module of the reference's
lib/modules/multiblock_cnn.py:92-335 (same constructor and forward signature).

Each block's input is padded with the cells of the blocks across its connected and periodic faces (zeros across
fixed boundaries), then convolved without padding.  The padding regions (the 3^D neighbourhood of a block, corners
and edges included) are resolved once, at construction, with the reference's traversal: from the block, each axis
in order and each side is crossed; a connected or periodic face leads into the neighbour (its axes composed into
the running axis map) and the remaining axes are crossed from there; a fixed face gives zeros; a region that a
later path reaches through a block replaces zeros, one already taken by a block is kept
(multiblock_cnn.py:151-214).  Each block region is then the neighbour's tensor flipped, permuted into the block's
orientation and cut to the padding width (:216-262).  Instead of replaying those tensor operations on every call,
they are applied once to the neighbours' cell indices: each block gets one index map into the concatenation of
all blocks' cells (plus a zero cell), so a forward pass is one gather per block followed by the convolution, and
autograd differentiates the gather.

The neighbour's orientation follows the reference exactly: it is flipped along the axes the inverted axis map marks,
then permuted by the inverted map's axis order.  That is the true orientation whenever the axis permutation is its
own inverse (every 2-D connection, 3-D connections that keep or swap axes); for a 3-D connection whose axes map as a
3-cycle the reference permutes the other way round, which fails with a shape error when the tangential sizes differ
and otherwise pads from a differently oriented neighbour.  Both outcomes are kept here (a RuntimeError at
construction for the first).
"""
import torch

from . import PISOtorch as P


def _zyx(i, dims):
    """tensor dimension (zyx order) <-> grid axis (xyz order)"""
    return dims - 1 - i


class _AxisMap:
    """Axis map from the padded block's frame to a neighbour's: axis i -> (axis ax[i], flipped inv[i])."""

    __slots__ = ("ax", "inv")

    def __init__(self, ax, inv):
        self.ax, self.inv = list(ax), list(inv)

    @classmethod
    def identity(cls, dims):
        return cls(range(dims), [False] * dims)

    @classmethod
    def across(cls, face, bound_axes, dims):
        """The map across a connected face (AxesMapping.from_bound, multiblock_cnn.py:55-62)."""
        a, up = face >> 1, face & 1
        ax, inv = [0] * dims, [False] * dims
        ax[a], inv[a] = bound_axes[0] >> 1, up == (bound_axes[0] & 1)
        for i in range(1, dims):
            ax[(a + i) % dims], inv[(a + i) % dims] = bound_axes[i] >> 1, bool(bound_axes[i] & 1)
        return cls(ax, inv)

    def then(self, other):
        """This map followed by `other` (make_fused_with :70-77)."""
        return _AxisMap([other.ax[t] for t in self.ax],
                        [other.inv[t] != v for t, v in zip(self.ax, self.inv)])

    def face(self, face):
        """A face of the padded block's frame in the neighbour's frame (transform_direction :79-82)."""
        a = face >> 1
        return self.ax[a] * 2 + int(bool(face & 1) != self.inv[a])

    def inverse(self):
        ax, inv = [0] * len(self.ax), [False] * len(self.ax)
        for i, (t, v) in enumerate(zip(self.ax, self.inv)):
            ax[t], inv[t] = i, v
        return _AxisMap(ax, inv)

    def permute_flip(self):
        """torch.flip dimensions, then torch.permute order, of a [B, C, (z), y, x] tensor (get_permute_flip_args
        :85-90)."""
        dims = len(self.ax)
        perm = [0, 1] + [_zyx(self.ax[_zyx(d, dims)], dims) + 2 for d in range(dims)]
        flip = [_zyx(d, dims) + 2 for d in range(dims) if self.inv[d]]
        return perm, flip


def _shape_zyx(block):
    dims = block.getSpatialDims()
    s = block.getSizes()
    return [s[_zyx(d, dims)] for d in range(dims)]


def _regions(block, blocks, dims):
    """{region: ("zero",) | ("block", index, axis map)}, region = per axis -1 / 0 / +1 (xyz order)."""
    info = {}

    def visit(blk, region, amap, axes):
        for a in axes:
            rest = [x for x in axes if x != a]
            for up in (0, 1):
                r = list(region)
                r[a] += 1 if up else -1
                r = tuple(r)
                face = amap.face(2 * a + up)
                bd = blk.getBoundaries()[face]
                if bd.type == P.BoundaryType.FIXED:
                    info.setdefault(r, ("zero",))
                elif bd.type in (P.BoundaryType.CONNECTED, P.BoundaryType.PERIODIC):
                    if r in info and info[r][0] != "zero":
                        continue
                    if bd.type == P.BoundaryType.CONNECTED:
                        nb = bd.getConnectedBlock()
                        nmap = amap.then(_AxisMap.across(face, list(bd.axes), dims))
                    else:
                        nb, nmap = blk, amap
                    info[r] = ("block", blocks.index(nb), nmap)
                    visit(nb, r, nmap, rest)
                else:
                    raise TypeError("Boundary type not supported.")

    visit(block, (0,) * dims, _AxisMap.identity(dims), list(range(dims)))
    return info


class MultiblockConv(torch.nn.Module):
    """A convolution over the blocks of a domain, each block padded from its neighbours
    (multiblock_cnn.py:92-335).  `fixed_boundary_padding`: "ZERO" or one of "VEL_U", "VEL_V", "VEL_W", "S", "ZERO"
    per input channel (validated; fixed boundaries pad with zeros, as in the reference).  kernel_size 3, 5 or 7,
    padding 'same' (= kernel_size // 2), dilation 1.  forward(tensor_list, domain=None) takes one [B, C, (z), y, x]
    tensor per block and returns one per block."""

    def __init__(self, domain, fixed_boundary_padding, in_channels, out_channels, kernel_size, stride=1,
                 padding="same", dilation=1, groups=1, bias=True, padding_mode="zeros", device=None, dtype=None):
        super().__init__()
        dims = domain.getSpatialDims()
        conv = {1: torch.nn.Conv1d, 2: torch.nn.Conv2d, 3: torch.nn.Conv3d}.get(dims)
        if conv is None:
            raise ValueError
        self.dims = dims
        per_channel = (isinstance(fixed_boundary_padding, (list, tuple)) and len(fixed_boundary_padding) == in_channels
                       and all(p in ("VEL_U", "VEL_V", "VEL_W", "S", "ZERO") for p in fixed_boundary_padding))
        if not (per_channel or fixed_boundary_padding == "ZERO"):
            raise ValueError
        self.fixed_boundary_padding = fixed_boundary_padding
        self.in_channels = in_channels
        if kernel_size not in (3, 5, 7):
            raise NotImplementedError("kernel_size must be 3, 5, or 7.")
        self.kernel_size = kernel_size
        if padding not in (kernel_size // 2, "same"):
            raise NotImplementedError("padding must be 1 or 'same'")
        self.padding = kernel_size // 2
        if dilation != 1:
            raise NotImplementedError("dilation is not supported")
        self.device, self.dtype = device, dtype
        blocks = domain.getBlocks()
        self.block_shapes = [_shape_zyx(b) for b in blocks]
        offsets, total = [], 0
        for s in self.block_shapes:
            offsets.append(total)
            n = 1
            for v in s:
                n *= v
            total += n
        self.total_cells = total
        for i, b in enumerate(blocks):
            self.register_buffer("_map%d" % i, self._index_map(i, b, blocks, offsets).to(device=device),
                                 persistent=False)
        self.conv_layer = conv(in_channels, out_channels, kernel_size, stride=stride, padding=0, dilation=dilation,
                               groups=groups, bias=bias, padding_mode=padding_mode, device=device, dtype=dtype)

    def _index_map(self, bi, block, blocks, offsets):
        """Indices of block bi's padded grid into [every block's cells (block order, zyx flattened), zero cell]."""
        dims, p = self.dims, self.padding
        info = _regions(block, blocks, dims)
        shape = self.block_shapes[bi]
        out = torch.full([s + 2 * p for s in shape], self.total_cells, dtype=torch.long)
        for idx in range(3 ** dims):
            region = tuple((idx // 3 ** a) % 3 - 1 for a in range(dims))  # xyz order
            dst = []
            for d in range(dims):  # zyx order
                r = region[_zyx(d, dims)]
                dst.append(slice(0, p) if r < 0 else slice(p + shape[d], 2 * p + shape[d]) if r > 0 else
                           slice(p, p + shape[d]))
            if all(r == 0 for r in region):
                src = torch.arange(offsets[bi], offsets[bi] + out[tuple(dst)].numel()).view(shape)
            else:
                entry = info.get(region, ("zero",))
                if entry[0] == "zero":
                    continue
                _, ni, amap = entry
                perm, flip = amap.inverse().permute_flip()
                nshape = self.block_shapes[ni]
                t = torch.arange(offsets[ni], offsets[ni] + int(torch.tensor(nshape).prod())).view([1, 1] + nshape)
                if flip:
                    t = torch.flip(t, flip)
                t = t.permute(perm)
                cut = []
                for d in range(dims):
                    r, m = region[_zyx(d, dims)], t.shape[d + 2]
                    cut.append(slice(m - p, m) if r < 0 else slice(0, p) if r > 0 else slice(None))
                src = t[(0, 0) + tuple(cut)]
            out[tuple(dst)] = src
        return out.reshape(-1)

    def _pad_all(self, tensor_list):
        B = tensor_list[0].size(0)
        flat = torch.cat([t.reshape(B, self.in_channels, -1) for t in tensor_list]
                         + [tensor_list[0].new_zeros(B, self.in_channels, 1)], dim=2)
        p = self.padding
        return [flat.index_select(2, getattr(self, "_map%d" % i)).view([B, self.in_channels]
                                                                       + [s + 2 * p for s in shape])
                for i, shape in enumerate(self.block_shapes)]

    def forward(self, tensor_list, domain=None):
        if domain is not None:
            if domain.getSpatialDims() != self.dims:
                raise ValueError("domain dimensionality does not match this layer.")
            if not (isinstance(tensor_list, (list, tuple)) and all(isinstance(t, torch.Tensor) for t in tensor_list)):
                raise TypeError("tensor_list must be a list of torch.Tensor.")
            if len(tensor_list) != domain.getNumBlocks():
                raise ValueError("length of tensor_list must match domain blocks.")
            if not tensor_list:
                raise ValueError("no input tensors.")
            if any(t.dim() != self.dims + 2 for t in tensor_list):
                raise ValueError("tensors must have shape NCDHW.")
            if any(t.size(0) != tensor_list[0].size(0) for t in tensor_list):
                raise ValueError("tensors have inconsistent batch size.")
            if any(t.size(1) != self.in_channels for t in tensor_list):
                raise ValueError("channel dimension of tensors must match in_channels.")
            if any(list(t.shape[2:]) != s for t, s in zip(tensor_list, self.block_shapes)):
                raise ValueError("tensor spatial dimensions must match block sizes")
        return [self.conv_layer(t) for t in self._pad_all(tensor_list)]

    def pad(self, tensor_list):
        """The padded input of every block (what the convolution sees)."""
        return self._pad_all(tensor_list)


class MultiblockConv_naive(torch.nn.Module):
    """One convolution applied to each block on its own, with the convolution's own padding
    (multiblock_cnn.py:337-357)."""

    def __init__(self, dims, *conv_args, **conv_kwargs):
        super().__init__()
        conv = {1: torch.nn.Conv1d, 2: torch.nn.Conv2d, 3: torch.nn.Conv3d}.get(dims)
        if conv is None:
            raise ValueError
        self.conv_layer = conv(*conv_args, **conv_kwargs)

    def forward(self, tensor_list, domain=None):
        return [self.conv_layer(t) for t in tensor_list]


class MultiblockReLU(torch.nn.Module):
    """ReLU of every block's tensor (multiblock_cnn.py:359-369)."""

    def forward(self, tensor_list, domain=None):
        return [torch.nn.functional.relu(t) for t in tensor_list]
