"""Model3D mirror (simulation/model_3d.{h,cpp}) for one partition.

``Step`` runs one cycle of ``UpdateEveryNthStep`` substeps on the device
(model_3d.cpp:35-51, without a host synchronization per substep), then
copies the points and indenter accumulators back and evaluates the cycle
force (model_3d.cpp:53-60, gpu_implementation6.cpp:417-510), and rebalances
(model_3d.cpp:76-108) when the free space or the share of disabled points
crosses its threshold: RemoveDisabledAndSort on the device, then the host
buffer refreshed from it. The reference's logging, timing log and snapshot
mutexes are not part of the step and are left out.
"""
from __future__ import annotations

from .capi import SimParams3D
from .gpu import GPUImplementation6


class Model3D:
    def __init__(self, prms: SimParams3D, device: int = 0):
        self.prms = prms
        self.gpu = GPUImplementation6(device)
        self.gpu.prms = prms
        self.force_history: list[tuple[float, float, float, float]] = []
        self.rebalances = 0

    def Reset(self) -> None:
        """model_3d.cpp:122-130: simulation time and step back to zero."""
        self.prms.SimulationStep = 0
        self.prms.SimulationTime = 0.0

    def Prepare(self) -> None:
        """model_3d.cpp:132-137."""
        self.gpu.update_constants(self.prms)

    def Step(self) -> bool:
        """One cycle; returns SimulationTime < SimulationEndTime (model_3d.cpp:111)."""
        p = self.prms
        self.gpu.reset_indenter_force_accumulator()
        n = p.UpdateEveryNthStep - (p.SimulationStep % p.UpdateEveryNthStep)
        t = self.gpu.step(n, p.SimulationTime, p.SimulationStep)
        p.SimulationTime = t
        p.SimulationStep += n
        self.gpu.transfer_from_device()
        f = self.gpu.indenter_force
        self.force_history.append((t, float((f ** 2).sum() ** 0.5), float(f[0]), float(f[1])))
        # rebalance (model_3d.cpp:76-108): too little free space or too many
        # disabled points -> squeeze and re-sort, here on the device
        hs = self.gpu.hssoa
        n = max(hs.size, 1)
        free_space = (hs.capacity - hs.size) / n
        disabled = hs.count_disabled() / max(hs.capacity, 1)
        if free_space < p.RebalanceThresholdFreeSpaceRemaining or disabled > p.RebalanceThresholdDisabledPercentage:
            self.gpu.remove_disabled_and_sort()
            # the reference sorts the host buffer and uploads it; here the
            # device sorted, and the host buffer is refreshed from it
            self.gpu.transfer_from_device()
            self.rebalances += 1
        return p.SimulationTime < p.SimulationEndTime
