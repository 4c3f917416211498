"""Backend selectors, the counterpart of src/backends.jl.

The reference dispatches `simulate_ensemble_bulk` on singleton backend types (backends.jl:14-44):
CPUBackend and GPUBackend for the Cartesian step, CPU_GCBackend and GPU_GCBackend for the
general-coordinate (tau, r) step.  The MI355X backend adds one type for each.
"""


class AbstractBackend:
    """backends.jl:14"""


class HIPBackend(AbstractBackend):
    """The Cartesian step on an MI355X (the role of GPUBackend, backends.jl:28)."""

    def __init__(self, device: int = 0):
        self.device = int(device)

    def __repr__(self):
        return f"HIPBackend(device={self.device})"


class HIP_GCBackend(AbstractBackend):
    """The general-coordinate (tau, r) step on an MI355X (the role of GPU_GCBackend,
    backends.jl:44, with the CPU_GCBackend signature of simulate.jl:117-133)."""

    def __init__(self, device: int = 0):
        self.device = int(device)

    def __repr__(self):
        return f"HIP_GCBackend(device={self.device})"
