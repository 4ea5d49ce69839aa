"""GpuState: the MI355X engine behind the reference's backend contract.

`GpuState.from_io_state` and `GpuState.produce_next_state` keep the names,
argument meaning and error split of rust/crates/gpu/src/gpu_state.rs:38-190
and :331-675: a fatal error raises (the outer Err), an error after which the
state is still valid comes back beside the state (the inner Result,
rust/crates/core/src/compute_thread.rs:165-169).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import abi
from .io_state import EngineOptions, FrameInput, GridNodes, InputConsts, IoState, Particles, make_consts


class GpuError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{message} (code {code})")
        self.code = code


@dataclass
class GpuRunParameters:
    """GpuRunParameters, rust/crates/gpu/src/gpu_state.rs:331-334, with the
    time-step fields of CpuRunParameters (rust/crates/cpu/src/cpu_state.rs:138-143):
    with adaptive_time_steps every substep takes the reference's two limits
    and steps by the allowed step of its AdaptiveTimeStepState, at most
    max_time_step (mpm_set_adaptive_time_steps); without it, max_time_step
    (when given) is the fixed step of the call."""

    target_time: float
    store_grid: bool = False
    max_time_step: Optional[float] = None
    adaptive_time_steps: bool = False


def _check(lib, rc: int):
    if rc < 0:
        raise GpuError(rc, lib.mpm_error_string(rc).decode())
    return rc


def available_gpus() -> list[str]:
    """GpuContext::available_gpus: HIP device ordinals as names."""
    lib = abi.load_library()
    return [f"hip:{i}" for i in range(lib.mpm_device_count())]


class GpuState:
    def __init__(self, lib, handle, n: int, consts: abi.MpmConsts):
        self._lib = lib
        self._h = handle
        self.n = n
        self._consts = consts
        self.last_status: Optional[abi.MpmStatus] = None

    # rust/crates/gpu/src/gpu_state.rs:38-190
    @classmethod
    def from_io_state(
        cls,
        gpu: int | str,
        consts: InputConsts,
        time_step: float,
        io_state: IoState,
        options: EngineOptions | None = None,
        tuning: dict | None = None,
    ) -> "GpuState":
        lib = abi.load_library()
        device = int(str(gpu).split(":")[-1]) if not isinstance(gpu, int) else gpu
        c = make_consts(consts, time_step, options, device)
        parts = io_state.particles.contiguous()
        view = parts.view()
        view.time = io_state.time
        t = abi.MpmTuning()
        for k, v in (tuning or {}).items():
            setattr(t, k, int(v))
        h = C.c_void_p()
        _check(lib, lib.mpm_create(C.byref(c), C.byref(view), C.byref(t), C.byref(h)))
        st = cls(lib, h, parts.count, c)
        st._parameters = parts.parameters.copy()
        st._initial_positions = parts.initial_positions.copy()
        st._with_jp = parts.plastic_jp is not None
        return st

    def close(self):
        if self._h:
            self._lib.mpm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def time(self) -> float:
        return self._lib.mpm_time(self._h)

    # rust/crates/gpu/src/gpu_state.rs:337-675
    def produce_next_state(
        self, frame_input: FrameInput, params: GpuRunParameters, cancel: Optional[C.c_int32] = None
    ) -> tuple[IoState, Optional[GpuError]]:
        if params.adaptive_time_steps:
            self.set_adaptive_time_steps(True, params.max_time_step if params.max_time_step is not None else float("inf"))
        else:
            self.set_adaptive_time_steps(False)
            if params.max_time_step is not None:
                self.set_time_step(params.max_time_step)
        fi, keep = frame_input.to_c()
        status = abi.MpmStatus()
        rc = self._lib.mpm_step_to(
            self._h, params.target_time, C.byref(fi), C.byref(cancel) if cancel is not None else None, C.byref(status)
        )
        self.last_status = status
        _check(self._lib, rc)
        state = self.to_io_state(params.store_grid)
        inner = GpuError(rc, self._lib.mpm_error_string(rc).decode()) if rc > 0 else None
        return state, inner

    def step_to(self, target_time: float, frame_input: FrameInput) -> abi.MpmStatus:
        """Advance without downloading the state (benchmarks)."""
        fi, keep = frame_input.to_c()
        status = abi.MpmStatus()
        rc = self._lib.mpm_step_to(self._h, target_time, C.byref(fi), None, C.byref(status))
        self.last_status = status
        _check(self._lib, rc)
        return status

    def to_io_state(self, store_grid: bool = False) -> IoState:
        p = Particles.empty(self.n, with_jp=self._with_jp)
        view = p.view()
        _check(self._lib, self._lib.mpm_read_particles(self._h, C.byref(view)))
        grid = self.read_grid() if store_grid else None
        return IoState(time=view.time, particles=p, grid_nodes=grid)

    def _read_grid(self, fn) -> GridNodes:
        cap = 1 << 16
        while True:
            ids = np.zeros((cap, 3), np.int32)
            bits = np.zeros(cap, np.uint32)
            m = np.zeros(cap, np.float32)
            v = np.zeros((cap, 3), np.float32)
            g = abi.MpmGridView()
            g.capacity = cap
            g.node_ids, g.collider_bits, g.masses, g.velocities = (abi.ptr(ids), abi.ptr(bits), abi.ptr(m), abi.ptr(v))
            rc = fn(self._h, C.byref(g))
            if rc == abi.ERR_CAPACITY:
                cap = int(g.count)
                continue
            _check(self._lib, rc)
            k = int(g.count)
            return GridNodes(ids[:k], bits[:k], m[:k], v[:k])

    def read_grid(self) -> GridNodes:
        return self._read_grid(self._lib.mpm_read_grid)

    def read_grid_momentum(self) -> GridNodes:
        return self._read_grid(self._lib.mpm_read_grid_momentum)

    def run_stage(self, stage: int, frame_input: FrameInput):
        fi, keep = frame_input.to_c()
        _check(self._lib, self._lib.mpm_run_stage(self._h, stage, C.byref(fi)))

    STAGES = ("keys", "sort", "block_list", "p2g", "grid_update", "g2p2g", "g2p", "copy_dead")

    def set_profiling(self, enabled: bool, stages=None):
        """HIP-event timing of every stage, or of the named `stages` only."""
        mask = 0
        if enabled:
            if stages is None:
                mask = 1
            else:
                mask = sum(1 << self.STAGES.index(s) for s in set(stages))
                if mask == 1:  # the keys stage alone: 1 would mean every stage
                    mask |= 1 << 31
        if mask >= 1 << 31:
            mask -= 1 << 32  # int32 argument
        _check(self._lib, self._lib.mpm_set_profiling(self._h, mask))

    def kernel_times(self, reset: bool = True) -> dict:
        cap = 16
        names = (C.c_char_p * cap)()
        ms = (C.c_double * cap)()
        launches = (C.c_uint64 * cap)()
        count = C.c_int32()
        _check(self._lib, self._lib.mpm_kernel_times(self._h, 1 if reset else 0, cap, names, ms, launches, C.byref(count)))
        return {names[i].decode(): (ms[i], int(launches[i])) for i in range(count.value)}

    @property
    def stream(self) -> int:
        return int(self._lib.mpm_stream(self._h) or 0)

    def set_closure(self, model, judge_stride: int = 0):
        """Learned snow stress (a closure.ClosureModel) in every substep, or the
        analytic stress again with model=None (mpm_set_closure)."""
        if model is None:
            _check(self._lib, self._lib.mpm_set_closure(self._h, None))
            return
        p = model.params
        if hasattr(p, "data_ptr"):  # a float32 device tensor: the trained parameters in place
            import torch

            if p.dtype != torch.float32 or not p.is_cuda or not p.is_contiguous():
                raise ValueError("closure parameters must be a contiguous float32 device tensor")
            torch.cuda.current_stream().synchronize()  # written on torch's stream
            params, size, ptr = p, p.numel(), p.data_ptr()
        else:
            params = np.ascontiguousarray(p, dtype=np.float32)
            size, ptr = params.size, params.ctypes.data
        if size != int(self._lib.mpm_mlp_param_count()):
            raise ValueError("closure parameters have the wrong length")
        c = abi.MpmClosure()
        c.params = ptr
        c.hull_lo[:] = [float(v) for v in model.hull_lo]
        c.hull_hi[:] = [float(v) for v in model.hull_hi]
        c.judge_stride = int(judge_stride)
        _check(self._lib, self._lib.mpm_set_closure(self._h, C.byref(c)))

    def time_step_limits(self) -> dict:
        """The reference's adaptive time-step limits over the current state
        (mpm_read_time_step_limits); +inf where a limit has no particle."""
        out = abi.MpmTimeStepLimits()
        _check(self._lib, self._lib.mpm_read_time_step_limits(self._h, C.byref(out)))
        return {k: float(getattr(out, k)) for k, _ in abi.MpmTimeStepLimits._fields_}

    def set_collider_mesh(self, mesh):
        """Mesh colliders (an io_state.ColliderMesh) from the next call on, or
        none with mesh=None (mpm_set_collider_mesh); the frame inputs then
        carry the mesh's vertex positions, frictions and dampings."""
        if mesh is None:
            _check(self._lib, self._lib.mpm_set_collider_mesh(self._h, None))
            return
        m, keep = mesh.to_c()
        _check(self._lib, self._lib.mpm_set_collider_mesh(self._h, C.byref(m)))

    def debug_cancel_after(self, substep: int):
        """Test hook: the following calls behave as cancelled before their
        substep `substep` (mpm_debug_cancel_after); negative turns it off."""
        _check(self._lib, self._lib.mpm_debug_cancel_after(self._h, int(substep)))

    def set_time_step(self, time_step: float):
        _check(self._lib, self._lib.mpm_set_time_step(self._h, float(time_step)))

    def set_adaptive_time_steps(self, adaptive: bool, max_time_step: float = float("inf")):
        """Per-substep adaptive steps (mpm_set_adaptive_time_steps)."""
        _check(self._lib, self._lib.mpm_set_adaptive_time_steps(self._h, 1 if adaptive else 0, float(max_time_step)))

    def adapt_time_step(self, max_time_step: float) -> float:
        """Sets the step of the following calls to the smallest of
        max_time_step and the current limits (allowed_time_step_without_prior,
        rust/crates/cpu/src/adaptive_time_step_state.rs:36-47, taken once per
        call instead of twice per substep); returns it."""
        # a limit without particles (None in the reference) counts as f32::MAX
        f32_max = float(np.finfo(np.float32).max)
        dt = float(np.float32(min([min(max_time_step, f32_max), *(min(v, f32_max) for v in self.time_step_limits().values())])))
        self.set_time_step(dt)
        return dt

    def closure_pairs(self, stride: int, phase: int = 0, capacity: int = 1 << 20):
        """Training pairs (x [k][6], t [k][3], float32 device tensors) of the
        live snow particles with id % stride == phase (mpm_closure_pairs)."""
        import torch

        dev = torch.device("cuda", torch.cuda.current_device())
        x = torch.empty((capacity, 6), dtype=torch.float32, device=dev)
        t = torch.empty((capacity, 3), dtype=torch.float32, device=dev)
        torch.cuda.current_stream().synchronize()  # the buffers exist before the engine's stream writes them
        count = C.c_uint64()
        rc = self._lib.mpm_closure_pairs(self._h, int(stride), int(phase), C.c_void_p(x.data_ptr()),
                                         C.c_void_p(t.data_ptr()), capacity, C.byref(count))
        if rc == abi.ERR_CAPACITY:
            return self.closure_pairs(stride, phase, int(count.value))
        _check(self._lib, rc)
        k = int(count.value)
        return x[:k], t[:k]

    def closure_stats(self, reset: bool = True) -> dict:
        """Network / fallback / judged counts and the judged relative stress error."""
        st = abi.MpmClosureStats()
        _check(self._lib, self._lib.mpm_read_closure_stats(self._h, 1 if reset else 0, C.byref(st)))
        rel = float(np.sqrt(st.err2 / st.ref2)) if st.ref2 > 0 else float("nan")
        return {"evaluated": int(st.evaluated), "fallbacks": int(st.fallbacks), "judged": int(st.judged),
                "relative_stress_error": rel}
