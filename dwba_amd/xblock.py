"""Host planning of the block-form MFMA contraction (dwba_xblock_* of
include/dwba_hip.h): per-grid factor tables, groups of up to 64 consecutive
pairs of a stacked alpha / beta matrix, their block ranges and the dispatch
order of the ranges.

The kernel cuts each grid into blocks of 8 points; a group's blocks are
split into ranges of `blocks_per_range` blocks at the same boundaries for
every group of a grid, one workgroup each, and a join adds the ranges in grid
order.  Ranges of one grid are dispatched to one XCD (workgroup i runs on XCD
i mod 8), range by range with the grid's groups side by side, so the groups
read each block of the grid's factor table from that XCD's L2 at about the
same time: grids are assigned to the eight dispatch queues longest first (LPT
on blocks x pairs) and the queues interleaved.
"""
from __future__ import annotations

import ctypes
from typing import List, Sequence

import numpy as np

MODES = {"fp32": 0, "bf16x3": 1}   # DWBA_XBLOCK_F32, DWBA_XBLOCK_BF16X3
BLOCK = 8                          # grid points per block
GROUP = 64                         # pairs per workgroup
N_XCD = 8


class XBlockGrid(ctypes.Structure):
    """dwba_xblock_grid of include/dwba_hip.h."""
    _fields_ = [("r", ctypes.c_void_p), ("tab", ctypes.c_void_p), ("N", ctypes.c_int32), ("pad_", ctypes.c_int32)]


class XBlockGroup(ctypes.Structure):
    """dwba_xblock_group of include/dwba_hip.h."""
    _fields_ = [("A", ctypes.c_void_p), ("B", ctypes.c_void_p), ("tab", ctypes.c_void_p), ("ld", ctypes.c_int32),
                ("n_pairs", ctypes.c_int32), ("n_blocks", ctypes.c_int32), ("n_ranges", ctypes.c_int32),
                ("range0", ctypes.c_int32), ("out_row0", ctypes.c_int32), ("slot", ctypes.c_int32),
                ("blocks_per_range", ctypes.c_int32)]


def build_tables(hip, lib, torch, dev, grids: Sequence, mode: str, stream=None) -> List:
    """grids: (r device tensor, N) per grid.  Returns the table tensors (one
    per grid, in order), written on `stream` by dwba_xblock_tables."""
    m = MODES[mode]
    tabs = []
    arr = (XBlockGrid * len(grids))()
    for g, (r_d, N) in enumerate(grids):
        nb = int(lib.dwba_xblock_table_bytes(int(N), m))
        if nb <= 0:
            raise ValueError(f"dwba_xblock_table_bytes({N}, {m}) = {nb}")
        t = torch.empty(nb, dtype=torch.uint8, device=dev)
        tabs.append(t)
        arr[g] = XBlockGrid(hip.ptr(r_d), hip.ptr(t), int(N), 0)
    d = hip.struct_array_to_device(arr, dev)
    hip.check(lib.dwba_xblock_tables(len(grids), hip.ptr(d), max(int(N) for _, N in grids), m,
                                     hip.stream_handle(stream)), "dwba_xblock_tables")
    return tabs, d


class XBlockPlan:
    """Groups, ranges and dispatch order for a set of stacked families.

    families: sequence of (grid index, A tensor, B tensor, ld, extents, out_row0,
    slot) with A / B of shape [P][ld] (pair c in row c) and extents[c] >= the
    alpha and beta supports of pair c."""

    def __init__(self, hip, lib, torch, dev, families: Sequence, tabs: Sequence, mode: str,
                 blocks_per_range: int = 64):
        self.blocks_per_range = int(blocks_per_range)
        per_block = int(lib.dwba_xblock_table_bytes(BLOCK, MODES[mode]))
        groups = []
        cost_of_grid = {}
        grid_of_group = []
        range0 = 0
        for (gi, A, B, ld, ext, row0, slot) in families:
            ext = np.asarray(ext, dtype=np.int64)
            P = len(ext)
            for c0 in range(0, P, GROUP):
                n = min(GROUP, P - c0)
                e = int(ext[c0:c0 + n].max()) if n else 0
                nb = (e + BLOCK - 1) // BLOCK
                if n == 0 or nb == 0:
                    continue
                if BLOCK * nb > ld or nb * per_block > tabs[gi].numel():
                    raise ValueError(f"extent {e} beyond the grid (ld {ld}, table of "
                                     f"{tabs[gi].numel() // per_block} blocks)")
                nr = max(1, -(-nb // self.blocks_per_range))
                es = A.element_size()
                groups.append(XBlockGroup(hip.ptr(A) + c0 * ld * es, hip.ptr(B) + c0 * ld * es, hip.ptr(tabs[gi]),
                                          int(ld), n, nb, nr, range0, int(row0) + c0, int(slot),
                                          self.blocks_per_range))
                grid_of_group.append(gi)
                cost_of_grid[gi] = cost_of_grid.get(gi, 0) + nb * ((n + 15) // 16)
                range0 += nr
        self.n_groups = len(groups)
        self.n_ranges = range0
        # dispatch: whole grids to one of N_XCD queues (LPT), queues interleaved
        queues = [[] for _ in range(N_XCD)]
        load = np.zeros(N_XCD)
        for gi in sorted(cost_of_grid, key=lambda g: -cost_of_grid[g]):
            x = int(np.argmin(load))
            load[x] += cost_of_grid[gi]
            mine = [k for k in range(len(groups)) if grid_of_group[k] == gi]
            for rg in range(max(groups[k].n_ranges for k in mine)):
                queues[x] += [(k, rg) for k in mine if rg < groups[k].n_ranges]
        depth = max((len(q) for q in queues), default=0)
        items = np.full((depth * N_XCD, 2), -1, dtype=np.int32)
        items[:, 1] = 0
        for x, q in enumerate(queues):
            for i, it in enumerate(q):
                items[i * N_XCD + x] = it
        # drop trailing empty items
        while len(items) and items[-1, 0] < 0:
            items = items[:-1]
        self.n_items = len(items)
        self.groups_d = hip.struct_array_to_device((XBlockGroup * max(1, len(groups)))(*groups), dev) \
            if groups else None
        self.items_d = torch.as_tensor(np.ascontiguousarray(items), device=dev) if len(items) else None
        self._groups_host = groups

    def partial(self, lib, torch, dev):
        n = int(lib.dwba_xblock_partial_doubles(self.n_ranges))
        return torch.empty(max(n, 1), dtype=torch.float64, device=dev)

    def launch(self, hip, lib, mode: str, L_max: int, partial, out, stream=None):
        if not self.n_groups:
            return
        hip.check(lib.dwba_xblock_contract(MODES[mode], self.n_groups, hip.ptr(self.groups_d), self.n_items,
                                           hip.ptr(self.items_d), int(L_max), hip.ptr(partial), hip.ptr(out),
                                           hip.stream_handle(stream)), "dwba_xblock_contract")
