"""ctypes binding of the qcs C ABI (include/qcs.h + include/qcs_ext.h).

The same class drives either shared library that exports the ABI:

* ``libqcs.so`` -- the product: host C++ plus HIP kernels for gfx950
  (``rqs-svg_amd/lib/libqcs.so``);
* ``libqcs_oracle.so`` -- the CPU restatement of the reference plain method, used only
  by tests and by the benchmark's CPU-baseline leg.

Gate calls follow the reference signature (src/qcs.cu:3383-3759, include/qcs.h:108-141):
``gate(name, targets, params=(), ctrls=(), negctrls=())``.  Every C call returns 1/0;
a 0 raises :class:`QcsError` carrying the message the library sent to its exception
callback.
"""
from __future__ import annotations

import ctypes
import time
import os
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
PKG = HERE.parent
REPO = PKG.parent
GPU_LIB = PKG / "lib" / "libqcs.so"
ORACLE_LIB = REPO / "oracle" / "_build" / "libqcs_oracle.so"

bit_num_t = ctypes.c_int16
bit_t = ctypes.c_uint8

# name -> number of double parameters (include/qcs.h:108-141)
GATE_PARAMS = {
    "h": 0, "x": 0, "y": 0, "z": 0, "s": 0, "sdg": 0, "t": 0, "tdg": 0, "sx": 0,
    "sxdg": 0, "id": 0, "swap": 0, "iswap": 0, "dcx": 0, "ecr": 0, "rccx": 0, "rcccx": 0,
    "rx": 1, "ry": 1, "rz": 1, "u1": 1, "p": 1, "rxx": 1, "ryy": 1, "rzz": 1, "rzx": 1,
    "u2": 2, "r": 2, "xx_plus_yy": 2, "xx_minus_yy": 2,
    "u3": 3, "u": 3, "u4": 4,
}
GATE_TARGETS = {
    "h": 1, "x": 1, "y": 1, "z": 1, "s": 1, "sdg": 1, "t": 1, "tdg": 1, "sx": 1,
    "sxdg": 1, "id": 1, "rx": 1, "ry": 1, "rz": 1, "u1": 1, "p": 1, "u2": 1, "r": 1,
    "u3": 1, "u": 1, "u4": 1, "swap": 2, "iswap": 2, "dcx": 2, "ecr": 2, "rxx": 2,
    "ryy": 2, "rzz": 2, "rzx": 2, "xx_plus_yy": 2, "xx_minus_yy": 2, "rccx": 3,
    "rcccx": 4, "global_phase": 0,
}

KERNEL_CLASSES = ["pair", "phase", "diag", "dense4", "mono", "reduce", "collapse",
                  "sample", "permute", "init", "fused", "cast"]


class QcsError(RuntimeError):
    pass


_CB_TYPE = ctypes.CFUNCTYPE(None, ctypes.c_char_p, ctypes.c_size_t)


class Library:
    """One loaded ABI library plus its exception-callback plumbing."""

    _cache: dict = {}

    def __new__(cls, path, mode=None):
        key = (str(path), mode)
        if key not in cls._cache:
            obj = super().__new__(cls)
            obj._init(path, mode)
            cls._cache[key] = obj
        return cls._cache[key]

    def _init(self, path, mode):
        path = pathlib.Path(path)
        if not path.exists():
            raise FileNotFoundError(f"{path} is not built; run __graft_entry__.build()")
        self.path = path
        self.lib = ctypes.CDLL(str(path), mode=mode if mode is not None else ctypes.RTLD_LOCAL)
        self.last_error = ""

        def _cb(msg, length):
            self.last_error = ctypes.string_at(msg, length).decode(errors="replace")

        self._cb = _CB_TYPE(_cb)
        self.lib.qcs_set_exception_callback(self._cb)
        P = ctypes.POINTER
        vp = ctypes.c_void_p
        lib = self.lib
        lib.qcs_simulator_create.argtypes = [P(vp)]
        for name in ["destroy", "allocate_memory", "dispose", "set_zero_state",
                     "set_sequential_state", "set_flat_state", "set_entangled_state",
                     "set_random_state", "reset_clbits", "reset_measurement_state",
                     "reinitialize_mapping"]:
            getattr(lib, "qcs_simulator_" + name).argtypes = [vp]
        for name in ["get_num_procs", "get_proc_num", "get_num_qubits", "get_num_clbits"]:
            getattr(lib, "qcs_simulator_" + name).argtypes = [vp, P(bit_t)]
        lib.qcs_simulator_set_num_qubits.argtypes = [vp, bit_num_t]
        lib.qcs_simulator_set_num_clbits.argtypes = [vp, bit_num_t]
        lib.qcs_simulator_set_mapping.argtypes = [vp, P(bit_num_t), bit_num_t]
        lib.qcs_simulator_get_clbits.argtypes = [vp, P(bit_t)]
        lib.qcs_simulator_get_clbits_string.argtypes = [vp, ctypes.c_char_p]
        lib.qcs_simulator_reset.argtypes = [vp, bit_num_t]
        lib.qcs_simulator_measure.argtypes = [vp, bit_num_t, P(bit_t)]
        lib.qcs_simulator_measure_to_clbit.argtypes = [vp, bit_num_t, bit_num_t, P(bit_t)]
        lib.qcs_simulator_read.argtypes = [vp, bit_num_t, P(bit_t)]
        lib.qcs_simulator_save_statevector.argtypes = [vp, ctypes.c_char_p]
        lib.qcs_simulator_event_create.argtypes = [vp, P(bit_t)]
        lib.qcs_simulator_event_record.argtypes = [vp, ctypes.c_int]
        lib.qcs_simulator_event_get_elapsed_time.argtypes = [vp, ctypes.c_int16, ctypes.c_int16, P(ctypes.c_double)]
        lists = [P(bit_num_t), bit_num_t] * 3
        for g, npar in GATE_PARAMS.items():
            getattr(lib, "qcs_simulator_gate_" + g).argtypes = [vp] + [ctypes.c_double] * npar + lists
        lib.qcs_simulator_gate_global_phase.argtypes = [vp, ctypes.c_double] + [P(bit_num_t), bit_num_t] * 2
        u64 = ctypes.c_uint64
        lib.qcs_ext_set_seed.argtypes = [vp, u64]
        lib.qcs_ext_get_statevector.argtypes = [vp, vp, u64]
        lib.qcs_ext_set_statevector.argtypes = [vp, vp, u64]
        lib.qcs_ext_load_statevector.argtypes = [vp, ctypes.c_char_p]
        lib.qcs_ext_sample.argtypes = [vp, u64, u64, vp]
        lib.qcs_ext_sample_noalloc.argtypes = [vp, u64, u64, vp]
        lib.qcs_ext_probability_total.argtypes = [vp, P(ctypes.c_double)]
        lib.qcs_ext_norm.argtypes = [vp, P(ctypes.c_double)]
        lib.qcs_ext_expect_z.argtypes = [vp, vp, ctypes.c_int]
        lib.qcs_ext_set_renormalize.argtypes = [vp, ctypes.c_int]
        lib.qcs_ext_set_local_qubits.argtypes = [vp, bit_num_t]
        lib.qcs_ext_get_mapping.argtypes = [vp, P(bit_num_t), bit_num_t]
        lib.qcs_ext_permute.argtypes = [vp, P(bit_num_t), bit_num_t]
        lib.qcs_ext_set_stage_bits.argtypes = [vp, ctypes.c_int]
        lib.qcs_ext_set_stage_stream.argtypes = [vp, ctypes.c_int]
        lib.qcs_ext_planner_set_check.argtypes = [vp, ctypes.c_int]
        lib.qcs_ext_set_precision.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        lib.qcs_ext_flush.argtypes = [vp]
        lib.qcs_ext_set_plan.argtypes = [vp] + [ctypes.c_int] * 6
        lib.qcs_ext_fusion_stats.argtypes = [vp, P(u64), P(u64)]
        lib.qcs_ext_overlap.argtypes = [vp, vp, vp]
        lib.qcs_ext_profile.argtypes = [vp, ctypes.c_int]
        lib.qcs_ext_profile_get.argtypes = [vp, ctypes.c_int, P(u64), P(ctypes.c_double), P(ctypes.c_double)]
        lib.qcs_ext_profile_reset.argtypes = [vp]
        f32p = P(ctypes.c_float)
        lib.qcs_ext_mlp_create.argtypes = [P(vp), ctypes.c_int, u64]
        lib.qcs_ext_mlp_destroy.argtypes = [vp]
        lib.qcs_ext_mlp_num_params.argtypes = [vp, P(ctypes.c_int)]
        lib.qcs_ext_mlp_get_params.argtypes = [vp, vp, ctypes.c_int]
        lib.qcs_ext_mlp_set_params.argtypes = [vp, vp, ctypes.c_int]
        lib.qcs_ext_mlp_predict.argtypes = [vp, vp, ctypes.c_int, vp]
        lib.qcs_ext_mlp_gradients.argtypes = [vp, vp, vp, ctypes.c_int, vp, f32p]
        lib.qcs_ext_mlp_train.argtypes = [vp, vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float, u64, f32p]
        lib.qcs_ext_planner_num_features.argtypes = [P(ctypes.c_int)]
        lib.qcs_ext_planner_record.argtypes = [vp, ctypes.c_int, u64, vp, vp]
        lib.qcs_ext_set_planner.argtypes = [vp, vp]
        lib.qcs_ext_planner_stats.argtypes = [vp, vp]
        lib.qcs_ext_planner_record_plans.argtypes = [vp, ctypes.c_int, vp, vp, vp]
        lib.qcs_ext_planner_last_choice.argtypes = [vp, vp]
        lib.qcs_ext_batch_create.argtypes = [P(vp)]
        lib.qcs_ext_batch_destroy.argtypes = [vp]
        lib.qcs_ext_batch_add_circuit.argtypes = [vp, bit_num_t, ctypes.c_int32, P(ctypes.c_char_p), P(bit_num_t), P(bit_num_t), P(ctypes.c_double)]
        lib.qcs_ext_batch_set_tuning.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        lib.qcs_ext_batch_set_precision.argtypes = [vp, ctypes.c_int]
        lib.qcs_ext_batch_autotune.argtypes = [vp, P(ctypes.c_int)]
        lib.qcs_ext_batch_run.argtypes = [vp, vp, u64, P(ctypes.c_double)]
        lib.qcs_ext_batch_run_plain.argtypes = [vp, vp, u64, P(ctypes.c_double)]
        lib.qcs_ext_batch_run_direct.argtypes = [vp, vp, u64, P(ctypes.c_double)]
        lib.qcs_ext_batch_info.argtypes = [vp, P(ctypes.c_int), P(ctypes.c_int), P(u64)]
        lib.qcs_ext_batch_traffic.argtypes = [vp, P(ctypes.c_double), P(u64)]

    def check(self, rc, what):
        if rc != 1:
            msg, self.last_error = self.last_error, ""
            raise QcsError(f"{what} failed: {msg}")


def gpu_library() -> Library:
    return Library(GPU_LIB)


def oracle_library() -> Library:
    return Library(ORACLE_LIB)


def _arr(values):
    values = [int(v) for v in values]
    if not values:
        return None, 0
    return (bit_num_t * len(values))(*values), len(values)


class Simulator:
    """Owns one qcs_simulator handle (create ... destroy)."""

    def __init__(self, library: Library, num_qubits: int, num_clbits: int = 0,
                 mapping=None, seed: int | None = None, precision: int = 0,
                 block_qubits: int = 5, allocate: bool = True, stage_bits: int | None = None):
        self.L = library
        self.lib = library.lib
        h = ctypes.c_void_p()
        self._call(self.lib.qcs_simulator_create(ctypes.byref(h)), "create")
        self.h = h
        self.n = num_qubits
        self._call(self.lib.qcs_simulator_set_num_qubits(self.h, num_qubits), "set_num_qubits")
        self._call(self.lib.qcs_simulator_set_num_clbits(self.h, num_clbits), "set_num_clbits")
        if mapping is not None:
            a, c = _arr(mapping)
            self._call(self.lib.qcs_simulator_set_mapping(self.h, a, c), "set_mapping")
        if seed is not None:
            self._call(self.lib.qcs_ext_set_seed(self.h, seed), "set_seed")
        if precision:
            self._call(self.lib.qcs_ext_set_precision(self.h, precision, block_qubits), "set_precision")
        if stage_bits is not None:
            self._call(self.lib.qcs_ext_set_stage_bits(self.h, stage_bits), "set_stage_bits")
        if allocate:
            self.allocate()

    def _call(self, rc, what):
        self.L.check(rc, what)

    def allocate(self):
        self._call(self.lib.qcs_simulator_allocate_memory(self.h), "allocate_memory")

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value is not None:
            self.lib.qcs_simulator_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # --- gates -------------------------------------------------------------------
    def gate(self, name, targets=(), params=(), ctrls=(), negctrls=()):
        n_ng, c_ng = _arr(negctrls)
        n_ps, c_ps = _arr(ctrls)
        if name == "global_phase":
            rc = self.lib.qcs_simulator_gate_global_phase(self.h, float(params[0]), n_ng, c_ng, n_ps, c_ps)
        else:
            t, ct = _arr(targets)
            fn = getattr(self.lib, "qcs_simulator_gate_" + name)
            rc = fn(self.h, *[float(p) for p in params], t, ct, n_ng, c_ng, n_ps, c_ps)
        self._call(rc, "gate_" + name)

    def run(self, circuit):
        """Apply a list of (name, targets, params, ctrls, negctrls) tuples."""
        for op in circuit:
            self.gate(*op)

    # --- measurement / state ----------------------------------------------------------
    def measure(self, q):
        r = bit_t()
        self._call(self.lib.qcs_simulator_measure(self.h, q, ctypes.byref(r)), "measure")
        return r.value

    def measure_to_clbit(self, q, c):
        r = bit_t()
        self._call(self.lib.qcs_simulator_measure_to_clbit(self.h, q, c, ctypes.byref(r)), "measure_to_clbit")
        return r.value

    def read(self, c):
        r = bit_t()
        self._call(self.lib.qcs_simulator_read(self.h, c, ctypes.byref(r)), "read")
        return r.value

    def reset(self, q):
        self._call(self.lib.qcs_simulator_reset(self.h, q), "reset")

    def clbits_string(self):
        nc = self.get("num_clbits")
        buf = ctypes.create_string_buffer(nc + 1)
        self._call(self.lib.qcs_simulator_get_clbits_string(self.h, buf), "get_clbits_string")
        return buf.value.decode()

    def clbits(self):
        nc = self.get("num_clbits")
        buf = (bit_t * max(nc, 1))()
        self._call(self.lib.qcs_simulator_get_clbits(self.h, buf), "get_clbits")
        return [buf[i] for i in range(nc)]

    def get(self, what):
        r = bit_t()
        self._call(getattr(self.lib, "qcs_simulator_get_" + what)(self.h, ctypes.byref(r)), "get_" + what)
        return r.value

    def call(self, what):
        self._call(getattr(self.lib, "qcs_simulator_" + what)(self.h), what)

    def statevector(self):
        out = np.empty(1 << self.n, dtype=np.complex128)
        self._call(self.lib.qcs_ext_get_statevector(self.h, out.ctypes.data, out.size), "get_statevector")
        return out

    def set_statevector(self, psi):
        psi = np.ascontiguousarray(psi, dtype=np.complex128)
        self._call(self.lib.qcs_ext_set_statevector(self.h, psi.ctypes.data, psi.size), "set_statevector")

    def load(self, path):
        self._call(self.lib.qcs_ext_load_statevector(self.h, str(path).encode()), "load_statevector")

    def save(self, path):
        self._call(self.lib.qcs_simulator_save_statevector(self.h, str(path).encode()), "save_statevector")

    def sample(self, shots, seed):
        out = np.empty(shots, dtype=np.uint64)
        self._call(self.lib.qcs_ext_sample(self.h, seed, shots, out.ctypes.data), "sample")
        return out

    def sample_noalloc(self, shots, seed):
        """qcs_ext_sample_noalloc: the same outcomes, with no allocation by the library."""
        out = np.empty(shots, dtype=np.uint64)
        self._call(self.lib.qcs_ext_sample_noalloc(self.h, seed, shots, out.ctypes.data), "sample_noalloc")
        return out

    def probability_total(self):
        r = ctypes.c_double()
        self._call(self.lib.qcs_ext_probability_total(self.h, ctypes.byref(r)), "probability_total")
        return r.value

    def norm(self):
        r = ctypes.c_double()
        self._call(self.lib.qcs_ext_norm(self.h, ctypes.byref(r)), "norm")
        return r.value

    def expect_z(self):
        out = np.empty(self.n, dtype=np.float64)
        self._call(self.lib.qcs_ext_expect_z(self.h, out.ctypes.data, self.n), "expect_z")
        return out

    def set_seed(self, seed):
        self._call(self.lib.qcs_ext_set_seed(self.h, seed), "set_seed")

    def set_renormalize(self, on):
        self._call(self.lib.qcs_ext_set_renormalize(self.h, int(bool(on))), "set_renormalize")

    def set_local_qubits(self, nl):
        self._call(self.lib.qcs_ext_set_local_qubits(self.h, nl), "set_local_qubits")

    def mapping(self):
        buf = (bit_num_t * self.n)()
        self._call(self.lib.qcs_ext_get_mapping(self.h, buf, self.n), "get_mapping")
        return [buf[i] for i in range(self.n)]

    def permute(self, p2l):
        a, c = _arr(p2l)
        self._call(self.lib.qcs_ext_permute(self.h, a, c), "permute")

    def set_stage_bits(self, m):
        self._call(self.lib.qcs_ext_set_stage_bits(self.h, m), "set_stage_bits")

    def set_stage_stream(self, on):
        """Stages on 2^13 tiles of 22-33 qubit states: streaming kernel (True) or one
        workgroup per tile (False); the same bits either way."""
        self._call(self.lib.qcs_ext_set_stage_stream(self.h, int(bool(on))), "set_stage_stream")

    def flush(self):
        self._call(self.lib.qcs_ext_flush(self.h), "flush")

    def set_plan(self, tile_bits=None, k=3, tiles=2, window=64, rule=0):
        """Fixed fusion plan for every flush (qcs_ext_set_plan); tile_bits None: default."""
        en = tile_bits is not None
        self._call(self.lib.qcs_ext_set_plan(self.h, int(en), tile_bits or 0, k, tiles, window, rule), "set_plan")

    def fusion_stats(self):
        g, b = ctypes.c_uint64(), ctypes.c_uint64()
        self._call(self.lib.qcs_ext_fusion_stats(self.h, ctypes.byref(g), ctypes.byref(b)), "fusion_stats")
        return g.value, b.value

    def overlap(self, other):
        """(<self|other>, max |self_i - other_i|); both simulators from the same library."""
        out = np.empty(3, dtype=np.float64)
        self._call(self.lib.qcs_ext_overlap(self.h, other.h, out.ctypes.data), "overlap")
        return complex(out[0], out[1]), float(out[2])

    # --- timing -----------------------------------------------------------------------
    def event_create(self):
        r = bit_t()
        self._call(self.lib.qcs_simulator_event_create(self.h, ctypes.byref(r)), "event_create")
        return r.value

    def event_record(self, e):
        self._call(self.lib.qcs_simulator_event_record(self.h, e), "event_record")

    def elapsed(self, a, b):
        r = ctypes.c_double()
        self._call(self.lib.qcs_simulator_event_get_elapsed_time(self.h, a, b, ctypes.byref(r)), "elapsed")
        return r.value

    def profile(self, on=True):
        self._call(self.lib.qcs_ext_profile(self.h, int(bool(on))), "profile")

    def profile_reset(self):
        self._call(self.lib.qcs_ext_profile_reset(self.h), "profile_reset")

    def profile_get(self):
        res = {}
        for i, name in enumerate(KERNEL_CLASSES):
            n = ctypes.c_uint64()
            s = ctypes.c_double()
            b = ctypes.c_double()
            self._call(self.lib.qcs_ext_profile_get(self.h, i, ctypes.byref(n), ctypes.byref(s), ctypes.byref(b)), "profile_get")
            if n.value:
                res[name] = {"launches": n.value, "seconds": s.value, "bytes": b.value}
        return res


class Mlp:
    """Cost-model MLP (F -> 64 -> 64 -> 1): GPU kernels in libqcs.so, CPU loops in the oracle."""

    def __init__(self, library: Library, num_features: int = 16, seed: int = 1):
        self.L = library
        self.lib = library.lib
        h = ctypes.c_void_p()
        self.L.check(self.lib.qcs_ext_mlp_create(ctypes.byref(h), num_features, seed), "mlp_create")
        self.h = h
        self.F = num_features
        c = ctypes.c_int()
        self.L.check(self.lib.qcs_ext_mlp_num_params(self.h, ctypes.byref(c)), "mlp_num_params")
        self.P = c.value

    def params(self):
        out = np.empty(self.P, dtype=np.float32)
        self.L.check(self.lib.qcs_ext_mlp_get_params(self.h, out.ctypes.data, self.P), "mlp_get_params")
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        self.L.check(self.lib.qcs_ext_mlp_set_params(self.h, p.ctypes.data, p.size), "mlp_set_params")

    def predict(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.empty(X.shape[0], dtype=np.float32)
        self.L.check(self.lib.qcs_ext_mlp_predict(self.h, X.ctypes.data, X.shape[0], y.ctypes.data), "mlp_predict")
        return y

    def gradients(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        g = np.empty(self.P, dtype=np.float32)
        loss = ctypes.c_float()
        self.L.check(self.lib.qcs_ext_mlp_gradients(self.h, X.ctypes.data, T.ctypes.data, X.shape[0], g.ctypes.data, ctypes.byref(loss)), "mlp_gradients")
        return g, loss.value

    def train(self, X, T, epochs=50, batch=256, lr=1e-3, seed=1):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        loss = ctypes.c_float()
        self.L.check(self.lib.qcs_ext_mlp_train(self.h, X.ctypes.data, T.ctypes.data, X.shape[0], epochs, batch, lr, seed, ctypes.byref(loss)), "mlp_train")
        return loss.value

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value is not None:
            self.lib.qcs_ext_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def planner_num_features(library):
    F = ctypes.c_int()
    library.check(library.lib.qcs_ext_planner_num_features(ctypes.byref(F)), "planner_num_features")
    return F.value


def planner_record(sim, count, seed):
    """Times `count` random launches (blocks, and stages on large enough states) on the
    simulator's state: (features [count, F], log2 microseconds [count])."""
    F = planner_num_features(sim.L)
    X = np.empty((count, F), dtype=np.float32)
    y = np.empty(count, dtype=np.float32)
    sim._call(sim.lib.qcs_ext_planner_record(sim.h, count, seed, X.ctypes.data, y.ctypes.data), "planner_record")
    return X, y


def planner_record_plans(sim, capacity=4096):
    """Times the launches of every stage plan the planner considers for the gates queued on
    `sim` (applied once per plan): (features [count, F], log2 microseconds [count])."""
    F = planner_num_features(sim.L)
    X = np.empty((capacity, F), dtype=np.float32)
    y = np.empty(capacity, dtype=np.float32)
    c = ctypes.c_int()
    sim._call(sim.lib.qcs_ext_planner_record_plans(sim.h, capacity, X.ctypes.data, y.ctypes.data, ctypes.byref(c)), "planner_record_plans")
    return X[: c.value].copy(), y[: c.value].copy()


def planner_last_choice(sim):
    """(stage tile bits, block qubits, tile choice, window, rule) of the last learned flush."""
    out = (ctypes.c_int * 5)()
    sim._call(sim.lib.qcs_ext_planner_last_choice(sim.h, out), "planner_last_choice")
    return tuple(out)


def train_cost_model(library, precision, block_qubits, random_sizes, random_count, plan_sizes,
                     plan_seeds=(1001, 1002, 1003), epochs=300, seed=5, log=None):
    """Records launch timings on the GPU and trains a cost model on them with the GPU
    kernels: `random_count` random blocks and stages spread over `random_sizes` qubits
    (halving per size step), plus the launches of every stage plan the planner considers
    for the circuits of circuits.planner_training_set(m, s) at each m in `plan_sizes` and s
    in `plan_seeds`.  Returns (model, info)."""
    from . import circuits

    F = planner_num_features(library)
    Xs, ys = [], []
    t0 = time.perf_counter()
    weights = [0.5 ** i for i in range(len(random_sizes))]
    for i, m in enumerate(random_sizes):
        cnt = max(1, int(random_count * weights[i] / sum(weights)))
        with Simulator(library, m, precision=precision, block_qubits=block_qubits) as s:
            X, y = planner_record(s, cnt, seed=77 + i)
        Xs.append(X)
        ys.append(y)
    n_random = sum(len(y) for y in ys)
    t_random = time.perf_counter() - t0
    t0 = time.perf_counter()
    for m in plan_sizes:
        with Simulator(library, m, precision=precision, block_qubits=block_qubits) as s:
            for sd in plan_seeds:
                for circ in circuits.planner_training_set(m, sd).values():
                    s.run(circ)
                    X, y = planner_record_plans(s, capacity=8192)
                    Xs.append(X)
                    ys.append(y)
    t_plans = time.perf_counter() - t0
    X = np.concatenate(Xs)
    y = np.concatenate(ys)
    model = Mlp(library, F, seed=seed)
    t0 = time.perf_counter()
    loss = model.train(X, y, epochs=epochs, batch=256, lr=2e-3, seed=2)
    info = {"random_rows": int(n_random), "plan_rows": int(len(y) - n_random), "record_random_s": t_random,
            "record_plans_s": t_plans, "train_s": time.perf_counter() - t0, "train_mse_log2us": float(loss),
            "target_variance": float(np.var(y)), "random_sizes": list(random_sizes), "plan_sizes": list(plan_sizes)}
    if log:
        log(f"cost model (precision {precision}): {info['random_rows']} random + {info['plan_rows']} plan launches "
            f"recorded in {t_random + t_plans:.1f}s, trained in {info['train_s']:.1f}s, mse {loss:.4f} (variance {np.var(y):.4f})")
    return model, info


def set_planner(sim, model):
    sim._call(sim.lib.qcs_ext_set_planner(sim.h, model.h if model is not None else None), "set_planner")


def planner_set_check(sim, on):
    """Measured check of the model's overrides (default on): per flush structure, the model's
    non-default choice runs timed once, the default plan runs timed once, and later flushes
    of that structure run the faster of the two.  Off: every flush runs the model's choice."""
    sim._call(sim.lib.qcs_ext_planner_set_check(sim.h, int(bool(on))), "planner_set_check")


def planner_stats(sim):
    out = np.zeros(5, dtype=np.float64)
    sim._call(sim.lib.qcs_ext_planner_stats(sim.h, out.ctypes.data), "planner_stats")
    return {"flushes": out[0], "learned_choices": out[1], "pred_greedy_us": out[2],
            "pred_chosen_us": out[3], "blocks": out[4]}


class Batch:
    """Batching form: independent circuits of 1-target gates (at most one control each).

    A circuit is a list of (name, target, params, control) with control -1 for none."""

    def __init__(self, library: Library):
        self.L = library
        self.lib = library.lib
        h = ctypes.c_void_p()
        self.L.check(self.lib.qcs_ext_batch_create(ctypes.byref(h)), "batch_create")
        self.h = h
        self.sizes = []

    def add(self, n, gates):
        k = len(gates)
        names = (ctypes.c_char_p * max(k, 1))(*[g[0].encode() for g in gates])
        tg = (bit_num_t * max(k, 1))(*[g[1] for g in gates])
        ct = (bit_num_t * max(k, 1))(*[g[3] for g in gates])
        flat = []
        for g in gates:
            p = list(g[2]) + [0.0] * (4 - len(g[2]))
            flat += p
        pa = (ctypes.c_double * max(4 * k, 1))(*flat)
        self.L.check(self.lib.qcs_ext_batch_add_circuit(self.h, n, k, names, tg, ct, pa), "batch_add_circuit")
        self.sizes.append(n)

    def set_tuning(self, lds_max_qubits=14, block_size=512, block_qubits=5, streams=4, stage_bits=12):
        self.L.check(self.lib.qcs_ext_batch_set_tuning(self.h, lds_max_qubits, block_size, block_qubits, streams, stage_bits), "batch_set_tuning")

    def set_precision(self, mode):
        """1: FP32 MFMA block products (default); 3: BF16x3 split-precision MFMA."""
        self.L.check(self.lib.qcs_ext_batch_set_precision(self.h, mode), "batch_set_precision")

    def autotune(self):
        t = (ctypes.c_int * 5)()
        self.L.check(self.lib.qcs_ext_batch_autotune(self.h, t), "batch_autotune")
        return {"lds_max_qubits": t[0], "block_size": t[1], "block_qubits": t[2], "streams": t[3],
                "stage_bits": t[4]}

    def run(self):
        """Returns (list of <Z> arrays per circuit, device seconds)."""
        out = np.empty(sum(self.sizes), dtype=np.float64)
        sec = ctypes.c_double()
        self.L.check(self.lib.qcs_ext_batch_run(self.h, out.ctypes.data, out.size, ctypes.byref(sec)), "batch_run")
        res, off = [], 0
        for n in self.sizes:
            res.append(out[off:off + n].copy())
            off += n
        return res, sec.value

    def run_direct(self):
        """The batch's launches without the graph; (results, device seconds)."""
        out = np.empty(sum(self.sizes), dtype=np.float64)
        sec = ctypes.c_double()
        self.L.check(self.lib.qcs_ext_batch_run_direct(self.h, out.ctypes.data, out.size, ctypes.byref(sec)), "batch_run_direct")
        res, off = [], 0
        for n in self.sizes:
            res.append(out[off:off + n].copy())
            off += n
        return res, sec.value

    def run_plain(self):
        """The plain method circuit after circuit (complex128, C++ loop); (results, wall seconds)."""
        out = np.empty(sum(self.sizes), dtype=np.float64)
        sec = ctypes.c_double()
        self.L.check(self.lib.qcs_ext_batch_run_plain(self.h, out.ctypes.data, out.size, ctypes.byref(sec)), "batch_run_plain")
        res, off = [], 0
        for n in self.sizes:
            res.append(out[off:off + n].copy())
            off += n
        return res, sec.value

    def info(self):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_uint64()
        self.L.check(self.lib.qcs_ext_batch_info(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)), "batch_info")
        return {"circuits": a.value, "total_qubits": b.value, "graph_nodes": c.value}

    def traffic(self):
        """Algorithmic HBM bytes and kernel launches of one run."""
        by, n = ctypes.c_double(), ctypes.c_uint64()
        self.L.check(self.lib.qcs_ext_batch_traffic(self.h, ctypes.byref(by), ctypes.byref(n)), "batch_traffic")
        return {"hbm_bytes": by.value, "kernel_launches": n.value}

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value is not None:
            self.lib.qcs_ext_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def load_statevector_file(path, n):
    data = np.fromfile(path, dtype=np.complex128)
    if data.size != (1 << n):
        raise ValueError(f"{path}: {data.size} amplitudes, expected {1 << n}")
    return data
