"""Python host mirror of libmvmc_hip.so (ctypes over include/mvmc_hip.h).

Names follow the reference's backend seams:
  pfaffian_inverse_batched  mvmc_gpu::pfaffian_inverse_batched / BatchedEngine::run_f64
                            (crates/mvmc-gpu/src/lib.rs:104-226)
  Ensemble.sample           SamplingEngine over W walkers (crates/mvmc-core/src/sampling/engine.rs:134-168)
  Ensemble.measure          MeasurementStages::measure_sample batched over saved samples
                            (crates/mvmc-core/src/stage_backend.rs:418-438)
  Ensemble.sr_cg            SrStages::cg_step + parameter update (crates/mvmc-core/src/sr_backend.rs:128-292)
  Ensemble.opt_step         one SR step of vmc_para_opt (crates/mvmc-core/src/run.rs:1294-1475)
There is no CPU fallback: a missing library or GPU raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("MVMC_HIP_LIB") or os.path.join(HERE, "_build", "libmvmc_hip.so")

_lib = None


class BackendUnavailable(RuntimeError):
    """mvmc_gpu::Error::BackendUnavailable (crates/mvmc-gpu/src/lib.rs:83-97)."""


class InvalidShape(ValueError):
    """mvmc_gpu::Error::InvalidShape."""


class BackendError(RuntimeError):
    """mvmc_gpu::Error::Backend / StageError::Failed."""


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise BackendUnavailable(f"libmvmc_hip.so not built: {LIB_PATH}")
        L = ctypes.CDLL(LIB_PATH)
        L.mvmc_last_error.restype = ctypes.c_char_p
        L.mvmc_pfinv_f64.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_void_p]
        L.mvmc_pfinv_f64_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                             ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.mvmc_pfinv_f64_inplace.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                             ctypes.c_void_p, ctypes.c_void_p]
        L.mvmc_pfinv_c128.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.mvmc_ens_create.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_longlong,
                                      ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        for name in ["mvmc_ens_destroy", "mvmc_ens_sample", "mvmc_ens_measure", "mvmc_ens_status"]:
            getattr(L, name).argtypes = [ctypes.c_void_p]
        L.mvmc_ens_sr_cg.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong)]
        L.mvmc_ens_opt_step.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.mvmc_sr_cg_solve.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 4 + \
            [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
             ctypes.POINTER(ctypes.c_longlong), ctypes.c_void_p]
        L.mvmc_sr_cg_apply.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 4 + \
            [ctypes.c_double, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]
        L.mvmc_sr_assemble_s_g.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                           ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong, ctypes.c_int,
                                           ctypes.c_double, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_void_p]
        L.mvmc_sr_gram_complex.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_void_p]
        L.mvmc_sr_s_g.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 5
        L.mvmc_group_create.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.POINTER(ctypes.c_void_p)]
        for name in ["mvmc_group_destroy", "mvmc_group_sample"]:
            getattr(L, name).argtypes = [ctypes.c_void_p]
        L.mvmc_group_run.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        L.mvmc_group_set_graph.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.mvmc_group_set_autotune.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.mvmc_group_opt_step.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.mvmc_group_timings.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.mvmc_ens_launch_log.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        L.mvmc_ens_write_opt_data.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        L.mvmc_ens_set_output_dir.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        L.mvmc_sr_gram_real.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.mvmc_sr_cholesky_solve.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.POINTER(ctypes.c_int), ctypes.c_void_p]
        _lib = L
    return _lib


EXPORTED_SYMBOLS = [
    "mvmc_last_error", "mvmc_abi_version", "mvmc_ens_launch_log", "mvmc_pfinv_f64_noalloc",
    "mvmc_sr_assemble_s_g", "mvmc_sr_gram_complex", "mvmc_sr_s_g", "mvmc_group_create", "mvmc_group_destroy",
    "mvmc_group_run", "mvmc_group_set_graph", "mvmc_group_set_autotune", "mvmc_group_opt_step", "mvmc_group_sample", "mvmc_group_timings", "mvmc_pfinv_f64", "mvmc_pfinv_f64_inplace", "mvmc_ens_create",
    "mvmc_ens_destroy", "mvmc_ens_dims", "mvmc_ens_get_params", "mvmc_ens_set_params", "mvmc_ens_sample",
    "mvmc_ens_measure", "mvmc_ens_sr_cg", "mvmc_ens_opt_step", "mvmc_ens_get_energy", "mvmc_ens_get_operands",
    "mvmc_ens_get_o_store", "mvmc_ens_get_walker", "mvmc_ens_get_tables", "mvmc_ens_get_config", "mvmc_ens_get_cg",
    "mvmc_ens_timings", "mvmc_ens_phase_clocks", "mvmc_ens_status", "mvmc_sr_cg_solve", "mvmc_ens_set_precision",
    "mvmc_ens_vjp", "mvmc_ens_vjp_adjoint", "mvmc_ens_log_psi", "mvmc_ens_sr_direct", "mvmc_pfinv_c128",
    "mvmc_ens_measure_configs", "mvmc_sr_cg_apply", "mvmc_sr_gram_real", "mvmc_sr_gram_real_f32", "mvmc_sr_cholesky_solve",
    "mvmc_ens_write_opt_data", "mvmc_ens_set_output_dir",
]


def _check(rc):
    if rc == 0:
        return
    msg = lib().mvmc_last_error().decode(errors="replace")
    if rc == -1:
        raise InvalidShape(msg)
    if rc == -2:
        raise BackendUnavailable(msg)
    raise BackendError(f"rc={rc}: {msg}")


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def pfaffian_inverse_batched(planes_dev, pf_dev, inv_dev, status_dev, n, count, stream=None):
    """Device-pointer form (torch CUDA tensors or raw ints): planes [count][n][n]
    column-major planes, outputs Pf, TRUE inverse, status."""
    def addr(x):
        return x if isinstance(x, int) else x.data_ptr()
    _check(lib().mvmc_pfinv_f64(ctypes.c_void_p(addr(planes_dev)), n, count, ctypes.c_void_p(addr(pf_dev)),
                                ctypes.c_void_p(addr(inv_dev)), ctypes.c_void_p(addr(status_dev)),
                                ctypes.c_void_p(stream or 0)))


def pfaffian_inverse_batched_noalloc(planes_dev, pf_dev, inv_dev, status_dev, n, count, stream=None):
    """mvmc_pfinv_f64_noalloc: pfaffian_inverse_batched without any device
    allocation (the work buffer is the library's per-device static array;
    n <= 1024; overlapping calls on one device race)."""
    def addr(x):
        return x if isinstance(x, int) else x.data_ptr()
    _check(lib().mvmc_pfinv_f64_noalloc(ctypes.c_void_p(addr(planes_dev)), n, count, ctypes.c_void_p(addr(pf_dev)),
                                        ctypes.c_void_p(addr(inv_dev)), ctypes.c_void_p(addr(status_dev)),
                                        ctypes.c_void_p(stream or 0)))


def pfaffian_inverse_batched_inplace(planes_dev, pf_dev, status_dev, n, count, stream=None):
    """In-place form (pfaffian_inverse_batch_in_place, crates/mvmc-core/src/stage_backend.rs:145-155):
    the planes are overwritten with the TRUE inverse (zero-filled where a
    plane fails)."""
    def addr(x):
        return x if isinstance(x, int) else x.data_ptr()
    _check(lib().mvmc_pfinv_f64_inplace(ctypes.c_void_p(addr(planes_dev)), n, count, ctypes.c_void_p(addr(pf_dev)),
                                        ctypes.c_void_p(addr(status_dev)), ctypes.c_void_p(stream or 0)))


def pfaffian_inverse_batched_c128(planes_dev, pf_dev, inv_dev, status_dev, n, count, order=0, stream=None):
    """Complex form (BatchedEngine::run_c64): complex128 device tensors (or raw
    addresses), planes [count][n][n] column-major; order 0 = zsktf2 (the
    batched reference), 1 = the C-order LTL of the production sampler."""
    def addr(x):
        return x if isinstance(x, int) else x.data_ptr()
    _check(lib().mvmc_pfinv_c128(ctypes.c_void_p(addr(planes_dev)), n, count, int(order),
                                 ctypes.c_void_p(addr(pf_dev)), ctypes.c_void_p(addr(inv_dev)),
                                 ctypes.c_void_p(addr(status_dev)), ctypes.c_void_p(stream or 0)))


class Ensemble:
    """W walkers of one Expert-mode model resident on the GPU."""

    def __init__(self, namelist, initial_def=None, walkers=1, seed=None, _handle=None):
        if _handle is None:
            h = ctypes.c_void_p()
            rc = lib().mvmc_ens_create(namelist.encode(), initial_def.encode() if initial_def else None, int(walkers),
                                       0 if seed is None else int(seed), 0 if seed is None else 1, ctypes.byref(h))
            _check(rc)
            self._owned = True
        else:
            h = _handle         # a run owned by an EnsembleGroup
            self._owned = False
        self.h = h
        self._destroy = lib().mvmc_ens_destroy   # kept: module globals may be gone at interpreter exit
        d = np.zeros(16, dtype=np.int32)
        lib().mvmc_ens_dims(self.h, _p(d))
        (self.nsite, self.ne, self.n, self.nqp, self.nproj, self.nslater, self.size, self.nsample, self.walkers,
         self.ex_path, self.interval, self.warmup, self.grid, self.lds_sampler, self.lds_measure,
         self.grid_measure) = [int(x) for x in d]

    def close(self):
        if getattr(self, "h", None):
            if getattr(self, "_owned", True):
                self._destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def sample(self):
        _check(lib().mvmc_ens_sample(self.h))

    def measure(self):
        _check(lib().mvmc_ens_measure(self.h))

    def set_output_dir(self, out_dir):
        """Per-step zvo_out / zvo_var output into out_dir (None: off)."""
        _check(lib().mvmc_ens_set_output_dir(self.h, None if out_dir is None else str(out_dir).encode()))

    def write_opt_data(self, out_dir):
        """Write {CParaFileHead}_opt.dat (+ block files) of the recorded window."""
        _check(lib().mvmc_ens_write_opt_data(self.h, str(out_dir).encode()))

    def measure_configs(self, ele_idx):
        """Local energy, O row and validity of each configuration (rows of
        ele_idx, EleIdx layout) with the current parameters."""
        cfg = np.ascontiguousarray(ele_idx, dtype=np.int32).reshape(-1, self.n)
        count = cfg.shape[0]
        e = np.zeros(count)
        o = np.zeros((count, self.size))
        ok = np.zeros(count, dtype=np.int32)
        _check(lib().mvmc_ens_measure_configs(self.h, ctypes.c_int(count), _p(cfg), _p(e), _p(o), _p(ok)))
        return e, o, ok

    def sr_cg(self):
        info = ctypes.c_longlong(0)
        _check(lib().mvmc_ens_sr_cg(self.h, ctypes.byref(info)))
        return info.value

    def sr_direct(self):
        """Direct SR (DSYRK + Cholesky) on the last measurement; 1 if skipped."""
        sk = ctypes.c_int(0)
        _check(lib().mvmc_ens_sr_direct(self.h, ctypes.byref(sk)))
        return sk.value

    def opt_step(self):
        row = np.zeros(6)
        _check(lib().mvmc_ens_opt_step(self.h, _p(row)))
        return row

    def params(self):
        p = np.zeros(max(self.nproj, 1))
        s = np.zeros(max(self.nslater, 1))
        lib().mvmc_ens_get_params(self.h, _p(p), _p(s))
        return p[:self.nproj].copy(), s[:self.nslater].copy()

    def vjp(self, cot):
        """sum_r cot[r] * O[r] over the measured samples (r = walker * NVMCSample
        + sample): [sum cot, sum cot * d log|psi| / d p]."""
        c = np.ascontiguousarray(cot, dtype=np.float64).reshape(-1)
        if c.size != self.walkers * self.nsample:
            raise InvalidShape(f"cotangent length {c.size} != walkers * NVMCSample")
        g = np.zeros(self.size)
        _check(lib().mvmc_ens_vjp(self.h, _p(c), _p(g)))
        return g

    def vjp_adjoint(self, cot):
        """The same VJP by the adjoint kernels (no O rows: inverse elements
        back onto the Slater table, the orbital map applied once); uses the
        saved configurations of the last sample call and rebuilds their
        planes."""
        c = np.ascontiguousarray(cot, dtype=np.float64).reshape(-1)
        if c.size != self.walkers * self.nsample:
            raise InvalidShape(f"cotangent length {c.size} != walkers * NVMCSample")
        g = np.zeros(self.size)
        _check(lib().mvmc_ens_vjp_adjoint(self.h, _p(c), _p(g)))
        return g

    def log_psi(self):
        """log|psi| of every saved sample at the current parameters, [W, NVMCSample]."""
        out = np.zeros(self.walkers * self.nsample)
        _check(lib().mvmc_ens_log_psi(self.h, _p(out)))
        return out.reshape(self.walkers, self.nsample)

    def set_precision(self, mode):
        """0: FP64 throughout; 1 ("mixed"): the sampler's resident inverse in
        FP32, FP64 arithmetic and everything else FP64."""
        _check(lib().mvmc_ens_set_precision(self.h, int({"fp64": 0, "mixed": 1}.get(mode, mode))))

    def set_params(self, proj, slater):
        p = np.ascontiguousarray(proj, dtype=np.float64)
        s = np.ascontiguousarray(slater, dtype=np.float64)
        lib().mvmc_ens_set_params(self.h, _p(p), _p(s))

    def energy(self):
        e = np.zeros(5)
        lib().mvmc_ens_get_energy(self.h, _p(e))
        return e

    def operands(self):
        oo = np.zeros(2 * self.size)
        ho = np.zeros(self.size)
        lib().mvmc_ens_get_operands(self.h, _p(oo), _p(ho))
        return oo, ho

    def o_store(self, walker=0):
        o = np.zeros(self.nsample * self.size)
        _check(lib().mvmc_ens_get_o_store(self.h, int(walker), _p(o)))
        return o.reshape(self.nsample, self.size)

    def walker(self, w):
        idx = np.zeros(self.nsample * self.n, dtype=np.int32)
        cnt = np.zeros(4, dtype=np.int64)
        rng = np.zeros(625, dtype=np.uint32)
        _check(lib().mvmc_ens_get_walker(self.h, int(w), _p(idx), _p(cnt), _p(rng)))
        return {"ele_idx": idx.reshape(self.nsample, self.n), "counter": cnt, "rng": rng[:624].copy(),
                "rng_idx": int(rng[624])}

    def tables(self, w):
        inv = np.zeros(self.nqp * self.n * self.n)
        pf = np.zeros(self.nqp)
        _check(lib().mvmc_ens_get_tables(self.h, int(w), _p(inv), _p(pf)))
        return inv.reshape(self.nqp, self.n, self.n), pf

    def cg(self):
        x = np.zeros(self.size)
        na = ctypes.c_int(0)
        st = lib().mvmc_ens_get_cg(self.h, _p(x), ctypes.byref(na))
        return st, x[:na.value].copy()

    def timings(self):
        t = np.zeros(26)
        lib().mvmc_ens_timings(self.h, _p(t))
        return {"sample_s": t[0], "measure_s": t[1], "sr_s": t[2], "proposals": t[3], "accepted": t[4],
                "block_ms": t[5], "block_launches": t[6], "rebuild_ms": t[7], "rebuilt": t[8],
                "meas_rebuild_ms": t[9], "measure_ms": t[10], "measured": t[11], "cg_ms": t[12],
                "cg_iter": t[13], "cg_na": t[14], "cg_rows": t[15], "rebuild_variant": int(t[16]),
                "rebuild_ms_per_walker": (t[17], t[18], t[19]), "adjoint_vjp_ms": t[20],
                "adjoint_vjp_rebuild_ms": t[21], "gram_ms": t[22], "block_write_passes": t[23],
                "block_flush_passes": t[24]}

    PHASES = ("candidate", "projection", "staging", "ratio", "metropolis", "accept_pass", "rebuild", "exchange",
              "block_end")
    REBUILD_PHASES = ("pf_pivot", "pf_scale_catchup", "pf_swap", "pf_record_eager", "pf_close", "inv_trtri",
                      "inv_solve", "inv_trmm_fill", "inv_perm", "pfaffian")

    def phase_clocks(self):
        """Per-phase device seconds of the sampling kernel and of the periodic
        rebuild's Pfaffian/inverse (summed over walkers); None unless the
        library was built with MVMC_PHASE_CLOCKS."""
        c = np.zeros(32, dtype=np.int64)
        if lib().mvmc_ens_phase_clocks(self.h, _p(c)) != 0:
            return None
        out = {k: c[i] / 1e8 for i, k in enumerate(self.PHASES)}
        out.update({k: c[16 + i] / 1e8 for i, k in enumerate(self.REBUILD_PHASES)})
        return out

    def launch_log(self):
        """Per k_block launch of the last sample call: rows of (device ms,
        proposals, accepted moves, walkers that advanced)."""
        n = lib().mvmc_ens_launch_log(self.h, None, 0)
        out = np.zeros((max(n, 1), 4))
        n = lib().mvmc_ens_launch_log(self.h, _p(out), n)
        return out[:n]

    def status(self):
        return int(lib().mvmc_ens_status(self.h))


class EnsembleGroup:
    """Independent runs (Expert-mode models of uneven Nsize) stepped together:
    one launch per kernel class over all runs' walkers (mvmc_group_*)."""

    def __init__(self, namelists, initial_defs=None, walkers=None):
        R = len(namelists)
        inits = initial_defs or [None] * R
        walkers = walkers or [1] * R
        nl = (ctypes.c_char_p * R)(*[x.encode() for x in namelists])
        ini = (ctypes.c_char_p * R)(*[x.encode() if x else None for x in inits])
        wk = np.ascontiguousarray(walkers, dtype=np.int32)
        h = ctypes.c_void_p()
        _check(lib().mvmc_group_create(R, nl, ini, _p(wk), ctypes.byref(h)))
        self.h = h
        self._destroy = lib().mvmc_group_destroy
        self.runs = []
        for r in range(R):
            e = ctypes.c_void_p()
            _check(lib().mvmc_group_run(self.h, r, ctypes.byref(e)))
            self.runs.append(Ensemble(None, _handle=e))

    def set_graph(self, rounds):
        _check(lib().mvmc_group_set_graph(self.h, int(rounds)))

    def set_autotune(self, on):
        _check(lib().mvmc_group_set_autotune(self.h, 1 if on else 0))

    def opt_step(self):
        rows = np.zeros((len(self.runs), 6))
        _check(lib().mvmc_group_opt_step(self.h, _p(rows)))
        return rows

    def sample(self):
        _check(lib().mvmc_group_sample(self.h))

    def timings(self):
        t = np.zeros(10)
        lib().mvmc_group_timings(self.h, _p(t))
        return {"sample_s": t[0], "measure_s": t[1], "sr_s": t[2], "round_ms": t[3], "rounds": int(t[4]),
                "graph_replays": int(t[5]), "grid_sampler": int(t[6]), "grid_rebuild": int(t[7]),
                "grid_measure": int(t[8]), "lds_sampler": int(t[9])}

    def close(self):
        if getattr(self, "h", None):
            for e in self.runs:
                e.h = None
            self._destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def sr_cg_apply(pack_dev, mean_dev, diag_dev, x_dev, z_dev, na, walkers, nsample, inv_w, shift, stream=None):
    """SrStages::cg_local_product: z = S x of the sampled operator on device
    operands (torch CUDA tensors)."""
    _check(lib().mvmc_sr_cg_apply(na, walkers, nsample, ctypes.c_void_p(pack_dev.data_ptr()),
                                  ctypes.c_void_p(mean_dev.data_ptr()), ctypes.c_void_p(diag_dev.data_ptr()),
                                  ctypes.c_void_p(x_dev.data_ptr()), inv_w, shift,
                                  ctypes.c_void_p(z_dev.data_ptr()), ctypes.c_void_p(stream or 0)))


def sr_assemble_s_g(oo_dev, ho_dev, map_dev, nm, ld, offset, sta_del, step_dt, s_dev, g_dev, oo_complex=False,
                    ho_complex=False, stream=None):
    """SrStages::assemble_s_g on device operands (torch CUDA tensors): S
    column-major [nm][nm] and g of the direct solve."""
    _check(lib().mvmc_sr_assemble_s_g(ctypes.c_void_p(oo_dev.data_ptr()), int(oo_complex),
                                      ctypes.c_void_p(ho_dev.data_ptr()), int(ho_complex),
                                      ctypes.c_void_p(map_dev.data_ptr()), nm, ld, offset, sta_del, step_dt,
                                      ctypes.c_void_p(s_dev.data_ptr()), ctypes.c_void_p(g_dev.data_ptr()),
                                      ctypes.c_void_p(stream or 0)))


def sr_gram_complex(store_dev, out_dev, n, samples, stream=None):
    """SrStages::gram_complex: out[i + j n] = sum_s O[i,s] conj(O[j,s]) of the
    complex [samples][n] store (torch complex128 CUDA tensors)."""
    _check(lib().mvmc_sr_gram_complex(n, samples, ctypes.c_void_p(store_dev.data_ptr()),
                                      ctypes.c_void_p(out_dev.data_ptr()), ctypes.c_void_p(stream or 0)))


def sr_s_g(o_dev, e_dev, w_dev, s_dev, g_dev, ns, np_, stream=None):
    """SrStages::sr_s_g: o column-major [ns][np] (torch CUDA tensors); S
    column-major [np][np] and g."""
    _check(lib().mvmc_sr_s_g(ctypes.c_void_p(o_dev.data_ptr()), ns, np_, ctypes.c_void_p(e_dev.data_ptr()),
                             ctypes.c_void_p(w_dev.data_ptr()), ctypes.c_void_p(s_dev.data_ptr()),
                             ctypes.c_void_p(g_dev.data_ptr()), ctypes.c_void_p(stream or 0)))


def sr_gram_real(store_dev, out_dev, n, samples, stream=None):
    """SrStages::gram_real: out = O O^T of the [n, samples] column-major store
    (torch CUDA tensors)."""
    _check(lib().mvmc_sr_gram_real(n, samples, ctypes.c_void_p(store_dev.data_ptr()),
                                   ctypes.c_void_p(out_dev.data_ptr()), ctypes.c_void_p(stream or 0)))


def sr_gram_real_f32(store_dev, out_dev, n, samples, stream=None):
    """Precision form of gram_real: float32 store, FP32-MFMA Gram with FP32
    accumulation, written as float64."""
    _check(lib().mvmc_sr_gram_real_f32(n, samples, ctypes.c_void_p(store_dev.data_ptr()),
                                       ctypes.c_void_p(out_dev.data_ptr()), ctypes.c_void_p(stream or 0)))


def sr_cholesky_solve(s_dev, rhs_dev, n, stream=None):
    """SrStages::cholesky_solve: rhs becomes the solution of S x = rhs; returns
    info (0 ok, k > 0 not positive definite, -1 non-finite solution)."""
    info = ctypes.c_int(0)
    _check(lib().mvmc_sr_cholesky_solve(n, ctypes.c_void_p(s_dev.data_ptr()), ctypes.c_void_p(rhs_dev.data_ptr()),
                                        ctypes.byref(info), ctypes.c_void_p(stream or 0)))
    return info.value


def sr_cg_solve(pack_dev, mean_dev, diag_dev, g_dev, x_dev, na, walkers, nsample, inv_w, shift, tol, max_it,
                variant=0, stream=None):
    """SrStages::cg_step on device operands (torch CUDA tensors)."""
    st = ctypes.c_longlong(0)
    _check(lib().mvmc_sr_cg_solve(na, walkers, nsample, ctypes.c_void_p(pack_dev.data_ptr()),
                                  ctypes.c_void_p(mean_dev.data_ptr()), ctypes.c_void_p(diag_dev.data_ptr()),
                                  ctypes.c_void_p(g_dev.data_ptr()), inv_w, shift, tol, max_it, variant,
                                  ctypes.c_void_p(x_dev.data_ptr()), ctypes.byref(st), ctypes.c_void_p(stream or 0)))
    return st.value
