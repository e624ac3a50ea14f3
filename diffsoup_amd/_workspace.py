"""Workspace management for the raster entry points.

The C-ABI never allocates: each call receives a caller-owned buffer sized by
``ds_raster_workspace_size`` that holds the per-(view, triangle) records, the
per-tile counters and offsets, the tile bins and an int64 stats block.  This
module keeps one such buffer per (device, stream) and sizes the bin capacity
from what previous calls produced, without synchronising the stream: the
stats block is copied to pinned host memory behind an event and read only
once that event has completed.  A batch that overflows the capacity still
renders exactly (the kernels fall back to walking every triangle of the
view); the next call simply gets a bigger workspace.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib

_DEFAULT_TILE = 16


class RasterWorkspace:
    def __init__(self, device: torch.device):
        self.device = device
        self.buf: torch.Tensor | None = None
        self.capacity = 0
        self.key = None
        self._host_stats = None
        self._event = None
        self._pairs_seen = 0

    def _absorb_stats(self) -> None:
        if self._event is not None and self._event.query():
            pairs = int(self._host_stats[_lib.DS_STAT_PAIRS])
            self._pairs_seen = max(self._pairs_seen, pairs)
            self._event = None

    def prepare(self, B: int, H: list[int], W: list[int], T: int, tile: int = _DEFAULT_TILE):
        """Return (buffer, nbytes, capacity, H_arr, W_arr) ready for a raster call."""
        self._absorb_stats()
        Ha, Wa = _lib.int_array(H), _lib.int_array(W)
        L = _lib.lib()
        n_tiles = int(L.ds_raster_tile_count(B, Ha, Wa, tile))
        if n_tiles < 0:
            raise _lib.DiffSoupError(L.ds_last_error().decode())
        want = max(4 * B * T, 2 * n_tiles, 1 << 16, int(1.25 * self._pairs_seen))
        key = (B, tuple(H), tuple(W), T, tile)
        if self.buf is None or key != self.key or want > self.capacity:
            cap = want if want > self.capacity or key != self.key else self.capacity
            nbytes = int(L.ds_raster_workspace_size(B, Ha, Wa, T, tile, cap))
            if nbytes == 0:
                raise _lib.DiffSoupError(L.ds_last_error().decode())
            if self.buf is None or self.buf.numel() < nbytes:
                self.buf = torch.empty(nbytes + 256, dtype=torch.uint8, device=self.device)
            self.capacity = cap
            self.key = key
        nbytes = int(L.ds_raster_workspace_size(B, Ha, Wa, T, tile, self.capacity))
        return self.aligned(), nbytes, self.capacity, Ha, Wa

    def aligned(self) -> int:
        base = self.buf.data_ptr()
        return (base + 255) & ~255

    def stats_tensor(self) -> torch.Tensor:
        """The int64 stats block as a device tensor view (DS_STAT_* indices)."""
        off = self.aligned() - self.buf.data_ptr()
        return self.buf[off:off + 8 * _lib.DS_STAT_COUNT].view(torch.int64)

    def note_call(self) -> None:
        """Queue a non-blocking copy of the stats block for the capacity heuristic."""
        if self._host_stats is None:
            self._host_stats = torch.zeros(_lib.DS_STAT_COUNT, dtype=torch.int64, pin_memory=True)
        self._host_stats.copy_(self.stats_tensor(), non_blocking=True)
        self._event = torch.cuda.Event()
        self._event.record(torch.cuda.current_stream(self.device))


_cache: dict = {}


def raster_workspace(device: torch.device) -> RasterWorkspace:
    key = (device.index if device.index is not None else torch.cuda.current_device(),
           torch.cuda.current_stream(device).cuda_stream)
    ws = _cache.get(key)
    if ws is None:
        ws = RasterWorkspace(device)
        _cache[key] = ws
    return ws
