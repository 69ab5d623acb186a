"""HostSideSOA mirror (simulation/host_side_soa.h:51-70).

A single float64 buffer of 28 lines x ``capacity`` points; line k starts at
``k * capacity`` (``getPointerToLine``, host_side_soa.h:65). Line 0 holds
int64 bit patterns (grain id, status bits) in double slots and is only ever
viewed as int64 here.
"""
from __future__ import annotations

import numpy as np

from . import capi


class HostSideSOA:
    def __init__(self, capacity: int = 0):
        self.Allocate(capacity)

    def Allocate(self, capacity: int) -> None:
        """HostSideSOA::Allocate (host_side_soa.cpp:70-85): zeroed buffer, size 0."""
        self.capacity = int(capacity)
        self.host_buffer = np.zeros(capacity * capi.N_PTS_ARRAYS, dtype=np.float64)
        self.size = 0

    def line(self, k: int) -> np.ndarray:
        return self.host_buffer[k * self.capacity:(k + 1) * self.capacity]

    def lines(self) -> np.ndarray:
        """(28, capacity) view."""
        return self.host_buffer.reshape(capi.N_PTS_ARRAYS, self.capacity)

    def utility(self) -> np.ndarray:
        return self.line(capi.IDX_UTILITY).view(np.int64)

    def pos(self) -> np.ndarray:
        return self.lines()[capi.POSX:capi.POSX + 3, :self.size]

    def vel(self) -> np.ndarray:
        return self.lines()[capi.VELX:capi.VELX + 3, :self.size]

    def Fe(self) -> np.ndarray:
        return self.lines()[capi.FE00:capi.FE00 + 9, :self.size]

    def Bp(self) -> np.ndarray:
        return self.lines()[capi.BP00:capi.BP00 + 9, :self.size]

    def InitializeBlock(self) -> None:
        """host_side_soa.cpp:88-97: Jp_inv = 1, Fe diagonal = 1 for points [0, size)."""
        L = self.lines()
        L[capi.IDX_JP_INV, :self.size] = 1.0
        for i in range(3):
            L[capi.FE00 + 4 * i, :self.size] = 1.0

    def copy(self) -> "HostSideSOA":
        other = HostSideSOA(0)
        other.capacity = self.capacity
        other.host_buffer = self.host_buffer.copy()
        other.size = self.size
        return other

    def count_active(self) -> int:
        return int(np.count_nonzero((self.utility()[:self.size] & capi.STATUS_DISABLED) == 0))

    def count_disabled(self) -> int:
        return self.size - self.count_active()
