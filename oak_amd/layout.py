"""State layout: OAK's memory layout of the model variables, the zone
partition and the permuted state vector, and the loading and saving of
ensembles from and to GHER files with the permutation on the GPU.

* ``MemLayout`` -- MemoryLayout_byfilenames (assimilation.F90:1620-1748):
  the variables one after the other (StartIndex, VarSize, varshape); a node
  is "sea" where its mask field differs from valex; the state keeps the sea
  nodes only (removeLandPoints): SeaIndex (full -> state, -1 for land) and
  invindex (state -> full).
* ``ZoneLayout`` -- initPartition (assimilation.F90:578-641) through
  oakg_partition: zone sizes, zoneIndex (permuted position -> state index)
  and invZoneIndex; ``gather_index`` is invindex(zoneIndex(l)), the full
  field index of permuted row l (loadVector, :869-880).
* ``load_ensemble`` / ``save_ensemble`` -- loadEnsemble (:1268-1359) and its
  inverse: member files read with ``oak_amd.ufile``, rows gathered into the
  permuted state (and scattered back, land set to valex) by
  oakg_gather_rows / oakg_scatter_rows on the GPU.
* ``state_index_map`` -- for a variable's grid, the column of H of each node
  (SeaIndex, then invZoneIndex: sub2indmv + the permutation of H's columns,
  :1781-1814, 2404-2409), the ``state`` array of oak_amd.assimilation's
  obs_operator grids.

All indices are 0-based."""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from .rrsqrt import default_context
from .ufile import uread, uwrite


@dataclass
class MemLayout:
    varshape: list          # per variable, its shape
    start: np.ndarray       # StartIndex (0-based) per variable in the full vector
    size: np.ndarray        # VarSize
    start_sea: np.ndarray   # StartIndexSea
    size_sea: np.ndarray    # VarSizeSea
    mask: np.ndarray        # [totsize] True = sea
    sea_index: np.ndarray   # [totsize] state index, -1 for land
    invindex: np.ndarray    # [effsize] full index of each state element

    @property
    def totsize(self):
        return len(self.mask)

    @property
    def effsize(self):
        return len(self.invindex)

    @property
    def nvar(self):
        return len(self.varshape)


def mem_layout(masks, valex=None):
    """MemoryLayout from one mask field per variable: an array (sea where it
    differs from its valex, or True where sea for a boolean array), or
    (array, valex) as uread returns it."""
    shapes, full = [], []
    for k, m in enumerate(masks):
        if isinstance(m, tuple):
            a, vx = m
            sea = np.asarray(a) != vx
        elif np.asarray(m).dtype == bool:
            a = np.asarray(m)
            sea = a
        else:
            a = np.asarray(m)
            sea = a != (valex[k] if isinstance(valex, (list, tuple)) else valex)
        shapes.append(a.shape)
        full.append(np.asarray(sea, bool).ravel(order="F"))
    size = np.array([len(f) for f in full], np.int64)
    start = np.concatenate([[0], np.cumsum(size)[:-1]]).astype(np.int64)
    size_sea = np.array([int(f.sum()) for f in full], np.int64)
    start_sea = np.concatenate([[0], np.cumsum(size_sea)[:-1]]).astype(np.int64)
    mask = np.concatenate(full) if full else np.zeros(0, bool)
    sea_index = np.full(len(mask), -1, np.int64)
    invindex = np.nonzero(mask)[0].astype(np.int64)
    sea_index[invindex] = np.arange(len(invindex))
    return MemLayout(shapes, start, size, start_sea, size_sea, mask, sea_index, invindex)


@dataclass
class ZoneLayout:
    zone_size: np.ndarray
    zone_index: np.ndarray      # permuted position -> state index
    inv_zone_index: np.ndarray  # state index -> permuted position

    @property
    def zone_ptr(self):
        return np.concatenate([[0], np.cumsum(self.zone_size)]).astype(np.int64)


def zone_layout(partition):
    """initPartition (assimilation.F90:578-641) of a partition vector over
    the state (one label per state element; gaps in the labels are allowed,
    :408-412)."""
    from .assimilation import partition as _part
    zs, zi, izi = _part(partition)
    return ZoneLayout(np.asarray(zs), np.asarray(zi), np.asarray(izi))


def gather_index(lay, zones):
    """invindex(zoneIndex(l)): the full field index of permuted row l."""
    return np.ascontiguousarray(lay.invindex[zones.zone_index], np.int64)


def state_index_map(lay, zones, v):
    """Column of H (permuted state index) of each node of variable v, -1 for
    land, in the variable's shape (Fortran order)."""
    s = lay.sea_index[lay.start[v]:lay.start[v] + lay.size[v]]
    out = np.where(s >= 0, zones.inv_zone_index[np.maximum(s, 0)], -1)
    return out.reshape(lay.varshape[v], order="F")


def _permute(gather, idx, full_len, src, dst, ctx):
    import torch
    L = _lib.load()
    ctx = ctx or default_context()
    n = len(idx)
    N = src.shape[1]
    elem = src.element_size()
    it = torch.as_tensor(idx, dtype=torch.int64, device=src.device)
    ctx.check(getattr(L, "oakg_gather_rows" if gather else "oakg_scatter_rows")(
        ctx.handle, n, N, it.data_ptr(), full_len, src.data_ptr(), src.stride(1),
        dst.data_ptr(), dst.stride(1), elem, _lib.OAKG_DEVICE_PTRS))
    return dst


def load_ensemble(files, lay, zones, device="cuda:0", dtype=None):
    """loadEnsemble: files[k][v] is member k's GHER file of variable v.
    Returns the permuted state ensemble (effsize x N, column-major) on the
    device."""
    import torch
    dtype = dtype or torch.float64
    N = len(files)
    full = torch.empty((N, lay.totsize), dtype=dtype, device=device).t()
    for k, member in enumerate(files):
        vec = np.concatenate([uread(f)[0].ravel(order="F") for f in member])
        full[:, k] = torch.as_tensor(vec).to(device, dtype)
    E = torch.empty((N, lay.effsize), dtype=dtype, device=device).t()
    return _permute(True, gather_index(lay, zones), lay.totsize, full, E, None)


def save_ensemble(Ea, files, lay, zones, valex=9999.0, prec=8):
    """The analysed ensemble back into GHER files (land set to valex)."""
    import torch
    N = Ea.shape[1]
    full = torch.full((N, lay.totsize), valex, dtype=Ea.dtype, device=Ea.device).t()
    _permute(False, gather_index(lay, zones), lay.totsize, Ea, full, None)
    host = full.cpu().numpy()
    for k, member in enumerate(files):
        for v, f in enumerate(member):
            a = host[lay.start[v]:lay.start[v] + lay.size[v], k]
            uwrite(f, a.reshape(lay.varshape[v], order="F"), valex=valex, prec=prec)
