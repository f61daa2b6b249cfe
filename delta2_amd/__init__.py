"""delta2_amd -- MI355X contact detection for Delta2 (HIP kernels behind a C ABI).

Python view of ``include/d2hip.h``.  The library ``libd2hip.so`` is built
in-tree by ``delta2_amd.build``; importing this package does not need a GPU,
but every compute call goes through the HIP library and fails loudly when it
is missing or no GPU is present -- there is no CPU fallback.

Reference-named helpers mirror the reference's collision-detection calls:

=============================================  =======================================================
``compare_trees_full``                          collision::compareTreesFull (+ filterContacts)
``compare_trees_full_continuous``               collision::compareTreesFullContinuous + step rule
``fine_collision_clusters_with_time_step_selection``  separate_clusters.cpp:462-599 + selectTimeStep clamp
=============================================  =======================================================
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libd2hip.so")

D2_OK, D2_ERR_ARG, D2_ERR_HIP, D2_ERR_CAPACITY, D2_ERR_STATE, D2_ERR_RUNTIME = range(6)
D2_FP64, D2_FP32 = 0, 1
BODY_STRIDE = 40
NODE_STRIDE = 8
BUCKET_STRIDE = 88

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_fp = ctypes.POINTER(ctypes.c_float)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_vp = ctypes.c_void_p
_lib = None


class D2Error(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"d2hip error {code}: {msg}")
        self.code = code


class CapacityError(D2Error):
    pass


class Tune(ctypes.Structure):
    """Launch shapes (d2_tune): blocks of the chunked list kernels, of the
    grid-stride kernels and of the per-pair post-processing."""
    _fields_ = [("chunk_blocks", ctypes.c_int), ("stream_blocks", ctypes.c_int), ("post_blocks", ctypes.c_int),
                ("gate_phases", ctypes.c_int)]

    def as_tuple(self):
        return (self.chunk_blocks, self.stream_blocks, self.post_blocks, self.gate_phases)


def lib():
    """Loads libd2hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m delta2_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "d2_ctx_create": ([ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)], ctypes.c_int),
        "d2_ctx_destroy": ([_vp], ctypes.c_int),
        "d2_last_error": ([_vp], ctypes.c_char_p),
        "d2_set_precision": ([_vp, ctypes.c_int], ctypes.c_int),
        "d2_set_capacity": ([_vp, ctypes.c_int64], ctypes.c_int),
        "d2_set_auto_grow": ([_vp, ctypes.c_int], ctypes.c_int),
        "d2_tree_build": ([_dp, ctypes.c_int, _ip, ctypes.c_int, _u64p, ctypes.c_int, _ip, _dp,
                           ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "d2_pcg_seed": ([ctypes.c_uint64, ctypes.c_uint64, _u64p], ctypes.c_int),
        "d2_trees_upload": ([_vp, ctypes.c_int, _ip, _ip, _dp], ctypes.c_int),
        "d2_bodies_upload": ([_vp, ctypes.c_int, _ip, _dp], ctypes.c_int),
        "d2_bodies_download": ([_vp, _dp], ctypes.c_int),
        "d2_project_future": ([_vp, _dp], ctypes.c_int),
        "d2_solver_bodies": ([_vp, ctypes.c_int, _dp, _dp], ctypes.c_int),
        "d2_solve": ([_vp, _dp, _ip, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "d2_solve_results": ([_vp, _dp], ctypes.c_int),
        "d2_commit_step": ([_vp], ctypes.c_int),
        "d2_solver_reset": ([_vp], ctypes.c_int),
        "d2_pairs_upload": ([_vp, ctypes.c_int, _ip, _dp], ctypes.c_int),
        "d2_narrow_run": ([_vp, ctypes.c_int, _i64p], ctypes.c_int),
        "d2_ccd_run": ([_vp, _i64p], ctypes.c_int),
        "d2_cluster_steps": ([_vp, _ip, ctypes.c_int, ctypes.c_float, _dp], ctypes.c_int),
        "d2_result_contacts": ([_vp, _ip, _ip, _dp], ctypes.c_int),
        "d2_result_pairs": ([_vp, _dp], ctypes.c_int),
        "d2_narrow": ([_vp, ctypes.c_int, _ip, ctypes.c_int, _i64p], ctypes.c_int),
        "d2_ccd": ([_vp, ctypes.c_int, _ip, _dp, _i64p], ctypes.c_int),
        "d2_set_profiling": ([_vp, ctypes.c_int], ctypes.c_int),
        "d2_num_kernels": ([], ctypes.c_int),
        "d2_kernel_name": ([ctypes.c_int], ctypes.c_char_p),
        "d2_kernel_stats": ([_vp, ctypes.c_int, _i64p, _dp, _dp], ctypes.c_int),
        "d2_reset_stats": ([_vp], ctypes.c_int),
        "d2_sync": ([_vp], ctypes.c_int),
        "d2_run_info": ([_vp, _i64p, ctypes.c_int], ctypes.c_int),
        "d2_mlp_num_params": ([], ctypes.c_int),
        "d2_mlp_set_params": ([_vp, _fp], ctypes.c_int),
        "d2_mlp_get_params": ([_vp, _fp], ctypes.c_int),
        "d2_mlp_infer": ([_vp, _fp, ctypes.c_int, _fp], ctypes.c_int),
        "d2_mlp_loss_grad": ([_vp, _fp, _fp, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float, _fp, _fp],
                             ctypes.c_int),
        "d2_mlp_train": ([_vp, _fp, _fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                          ctypes.c_float, ctypes.c_float, ctypes.c_uint64, _fp], ctypes.c_int),
        "d2_pair_features": ([_vp, _fp], ctypes.c_int),
        "d2_set_learned": ([_vp, ctypes.c_int, ctypes.c_float, ctypes.c_float], ctypes.c_int),
        "d2_set_tune": ([_vp, ctypes.POINTER(Tune)], ctypes.c_int),
        "d2_get_tune": ([_vp, ctypes.POINTER(Tune)], ctypes.c_int),
        "d2_set_graph": ([_vp, ctypes.c_int], ctypes.c_int),
        "d2_graph_stats": ([_vp, _i64p, _i64p], ctypes.c_int),
        "d2_autotune": ([_vp, ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Tune), _dp],
                        ctypes.c_int),
        "d2_clusters_upload": ([_vp, _ip, ctypes.c_int], ctypes.c_int),
        "d2_step_run": ([_vp, ctypes.c_float, ctypes.c_int, _dp, _i64p], ctypes.c_int),
        "d2_tree_bounds_upload": ([_vp, _fp, _dp], ctypes.c_int),
        "d2_tree_bounds_download": ([_vp, _fp, _dp], ctypes.c_int),
        "d2_bodies_last_upload": ([_vp, _dp], ctypes.c_int),
        "d2_bodies_groups_upload": ([_vp, _ip], ctypes.c_int),
        "d2_broad_phase": ([_vp, ctypes.c_int, ctypes.c_double, _i64p], ctypes.c_int),
        "d2_pairs_download": ([_vp, _ip, _dp], ctypes.c_int),
        "d2_clusters_build": ([_vp, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "d2_clusters_download": ([_vp, _ip, _ip, _ip, _dp, _dp], ctypes.c_int),
        "d2_set_touch_stats": ([_vp, ctypes.c_int], ctypes.c_int),
        "d2_set_pruning": ([_vp, ctypes.c_int], ctypes.c_int),
        "d2_touch_stats": ([_vp, _i64p, ctypes.c_int], ctypes.c_int),
        "d2_find_contacts_bucket": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip, _ip, ctypes.c_int64, _ip, _dp,
                                     _i64p], ctypes.c_int),
        "d2_find_contacts_bucket_noalloc": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip, _ip, ctypes.c_int64,
                                             _ip, _dp, _i64p], ctypes.c_int),
        "d2_filter_contacts": ([_vp, ctypes.c_int, _dp, ctypes.c_double, _ip, _ip, ctypes.POINTER(ctypes.c_int)],
                               ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def exported_symbols():
    """Names declared in include/d2hip.h (used by the export test)."""
    import re
    hdr = os.path.join(_HERE, "..", "include", "d2hip.h")
    with open(hdr) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(d2_[a-z_0-9]+)\s*\(", text)))


def _dptr(a):
    return a.ctypes.data_as(_dp)


def _iptr(a):
    return a.ctypes.data_as(_ip)


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def pcg_seed(initstate=42, initseq=54):
    """PCG32 state {state, inc} as pcg32_srandom_r(initstate, initseq)."""
    st = np.zeros(2, dtype=np.uint64)
    lib().d2_pcg_seed(initstate, initseq, st.ctypes.data_as(_u64p))
    return st


def tree_build(V, F, rng):
    """Surrogate tree of mesh (V, F) -- model::SurrogateTree(V, F, opt).

    ``rng`` is the shared PCG32 state (uint64[2]); it advances in place.
    Returns (nodes int32[n, 8], buckets float64[n, 88])."""
    V = _f64(V).reshape(-1, 3)
    F = _i32(F).reshape(-1, 3)
    assert rng.dtype == np.uint64 and rng.size == 2
    cap = max(2 * F.shape[0], 8)
    nodes = np.zeros((cap, NODE_STRIDE), dtype=np.int32)
    buckets = np.zeros((cap, BUCKET_STRIDE))
    n = ctypes.c_int(0)
    rc = lib().d2_tree_build(_dptr(V), V.shape[0], _iptr(F), F.shape[0], rng.ctypes.data_as(_u64p), cap,
                             _iptr(nodes), _dptr(buckets), ctypes.byref(n))
    if rc != D2_OK:
        raise D2Error(rc, "d2_tree_build failed")
    return nodes[: n.value].copy(), buckets[: n.value].copy()


class Result:
    """Contacts of one run in CSR order (offsets[n_pairs+1]) plus per-pair values."""

    def __init__(self, offsets, ints, reals, per_pair=None):
        self.offsets, self.ints, self.reals, self.per_pair = offsets, ints, reals, per_pair

    @property
    def n(self):
        return int(self.offsets[-1]) if self.offsets.size else 0

    def pair(self, p):
        s, e = self.offsets[p], self.offsets[p + 1]
        return self.ints[s:e], self.reals[s:e]


class Context:
    """One device context (a HIP stream and its buffers) -- include/d2hip.h."""

    def __init__(self, device=0, precision=D2_FP64):
        L = lib()
        h = _vp()
        rc = L.d2_ctx_create(device, precision, ctypes.byref(h))
        if rc != D2_OK:
            raise D2Error(rc, "d2_ctx_create failed (is a GPU visible?)")
        self.h = h
        self.n_pairs = 0
        self.n_clusters = 0
        self.n_out = 0

    def close(self):
        if getattr(self, "h", None):
            lib().d2_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc != D2_OK:
            msg = lib().d2_last_error(self.h)
            msg = msg.decode() if msg else ""
            cls = CapacityError if rc == D2_ERR_CAPACITY else D2Error
            raise cls(rc, msg)

    def set_precision(self, precision):
        self._check(lib().d2_set_precision(self.h, precision))

    def set_capacity(self, entries):
        self._check(lib().d2_set_capacity(self.h, int(entries)))

    def set_auto_grow(self, on=True):
        self._check(lib().d2_set_auto_grow(self.h, int(bool(on))))

    def upload_trees(self, trees):
        """trees: a list of (nodes, buckets), or flat arrays with attributes
        offs, nodes and buckets (scenes.FlatTrees; uploaded without a copy)."""
        if hasattr(trees, "offs"):
            offs, nodes, buckets = _i32(trees.offs), _i32(trees.nodes), _f64(trees.buckets)
        else:
            offs = np.zeros(len(trees) + 1, dtype=np.int32)
            offs[1:] = np.cumsum([t[0].shape[0] for t in trees])
            nodes = _i32(np.concatenate([t[0] for t in trees]))
            buckets = _f64(np.concatenate([t[1] for t in trees]))
        self._check(lib().d2_trees_upload(self.h, offs.size - 1, _iptr(offs), _iptr(nodes), _dptr(buckets)))

    def upload_bodies(self, tree_of_body, records):
        tob = _i32(tree_of_body)
        rec = _f64(records)
        self._check(lib().d2_bodies_upload(self.h, tob.size, _iptr(tob), _dptr(rec)))
        self.n_bodies = tob.size

    def download_bodies(self):
        out = np.zeros((self.n_bodies, BODY_STRIDE))
        self._check(lib().d2_bodies_download(self.h, _dptr(out)))
        return out

    def project_future(self, step_per_body):
        st = _f64(step_per_body)
        self._check(lib().d2_project_future(self.h, _dptr(st)))

    def upload_pairs(self, pairs, max_time=None):
        pairs = _i32(pairs).reshape(-1, 2)
        mt = None if max_time is None else _f64(max_time)
        self._check(lib().d2_pairs_upload(self.h, pairs.shape[0], _iptr(pairs), None if mt is None else _dptr(mt)))
        self.n_pairs = pairs.shape[0]

    def narrow_run(self, filter=True):
        n = ctypes.c_int64(0)
        self._check(lib().d2_narrow_run(self.h, int(bool(filter)), ctypes.byref(n)))
        self.n_out = n.value
        return n.value

    def ccd_run(self):
        n = ctypes.c_int64(0)
        self._check(lib().d2_ccd_run(self.h, ctypes.byref(n)))
        self.n_out = n.value
        return n.value

    def result(self, with_pairs=False):
        off = np.zeros(self.n_pairs + 1, dtype=np.int32)
        self._check(lib().d2_result_contacts(self.h, _iptr(off), None, None))
        n = int(off[-1])
        if n != self.n_out:
            raise D2Error(D2_ERR_STATE, f"CSR total {n} != run count {self.n_out}")
        ints = np.zeros((n, 8), dtype=np.int32)
        reals = np.zeros((n, 11))
        self._check(lib().d2_result_contacts(self.h, None, _iptr(ints), _dptr(reals)))
        per = None
        if with_pairs:
            per = np.zeros((self.n_pairs, 2))
            self._check(lib().d2_result_pairs(self.h, _dptr(per)))
        return Result(off, ints, reals, per)

    def upload_clusters(self, cluster_of_pair, n_clusters):
        cop = _i32(cluster_of_pair)
        self._check(lib().d2_clusters_upload(self.h, _iptr(cop), int(n_clusters)))
        self.n_clusters = int(n_clusters)

    def step_run(self, dt, filter=True, want_steps=False):
        """CCD + step rule + cluster steps + projection + discrete narrow phase
        (+ filter) over the resident pairs; returns the contact count, and the
        cluster steps when asked."""
        n = ctypes.c_int64(0)
        steps = np.zeros(self.n_clusters) if want_steps else None
        self._check(lib().d2_step_run(self.h, float(dt), int(bool(filter)), None if steps is None else _dptr(steps),
                                      ctypes.byref(n)))
        self.n_out = n.value
        return (n.value, steps) if want_steps else n.value

    def solver_bodies(self, mass, friction):
        """Per-body mass and friction coefficient for the solver."""
        m, f = _f64(mass), _f64(friction)
        self._check(lib().d2_solver_bodies(self.h, m.size, _dptr(m), _dptr(f)))

    def solve(self, gravity=(0.0, 0.0, -9.8)):
        """SequentialImpulses::solve over every cluster of the last filtered
        step run (the future states are replaced): (per-cluster status, 1
        solved / 0 refused; number refused)."""
        g = _f64(gravity)
        status = np.zeros(max(self.n_clusters, 1), dtype=np.int32)
        nf = ctypes.c_int(0)
        self._check(lib().d2_solve(self.h, _dptr(g), _iptr(status), ctypes.byref(nf)))
        return status[:self.n_clusters], nf.value

    def solve_results(self):
        """Per contact of the last step run's filtered result (its CSR order):
        normal, impulse, impulse offset, friction impulse, tangential velocity."""
        out = np.zeros((max(self.n_out, 1), 11))
        self._check(lib().d2_solve_results(self.h, _dptr(out)))
        return out[:self.n_out]

    def commit_step(self):
        """current <- future, last step <- the step, future <- extrapolate (dynamic bodies)."""
        self._check(lib().d2_commit_step(self.h))

    def solver_reset(self):
        """Empty the solver's warm store."""
        self._check(lib().d2_solver_reset(self.h))

    def set_graph(self, on=True):
        self._check(lib().d2_set_graph(self.h, int(bool(on))))

    def graph_stats(self):
        """(graph segments captured, graph launches) since the context was created."""
        c, n = ctypes.c_int64(0), ctypes.c_int64(0)
        self._check(lib().d2_graph_stats(self.h, ctypes.byref(c), ctypes.byref(n)))
        return c.value, n.value

    def set_tune(self, chunk_blocks, stream_blocks, post_blocks, gate_phases=3):
        t = Tune(int(chunk_blocks), int(stream_blocks), int(post_blocks), int(gate_phases))
        self._check(lib().d2_set_tune(self.h, ctypes.byref(t)))

    def autotune(self, kind="step", dt=None, filter=True, reps=3):
        """Picks the launch shapes by timing runs of `kind` ("narrow", "ccd",
        "step") on the resident inputs; returns (shape, ms per run)."""
        from .scenes import DT
        k = {"narrow": 0, "ccd": 1, "step": 2}[kind]
        t = Tune()
        ms = ctypes.c_double(0)
        self._check(lib().d2_autotune(self.h, k, float(DT if dt is None else dt), int(bool(filter)), int(reps),
                                      ctypes.byref(t), ctypes.byref(ms)))
        return t.as_tuple(), ms.value

    def get_tune(self):
        t = Tune()
        self._check(lib().d2_get_tune(self.h, ctypes.byref(t)))
        return t.as_tuple()

    def cluster_steps(self, cluster_of_pair, n_clusters, dt):
        cop = _i32(cluster_of_pair)
        out = np.zeros(n_clusters)
        self._check(lib().d2_cluster_steps(self.h, _iptr(cop), int(n_clusters), float(dt), _dptr(out)))
        return out

    def set_profiling(self, on=True):
        self._check(lib().d2_set_profiling(self.h, int(bool(on))))

    def reset_stats(self):
        self._check(lib().d2_reset_stats(self.h))

    def kernel_stats(self):
        L = lib()
        out = {}
        for i in range(L.d2_num_kernels()):
            n = ctypes.c_int64(0)
            ms = ctypes.c_double(0)
            items = ctypes.c_double(0)
            self._check(L.d2_kernel_stats(self.h, i, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(items)))
            out[L.d2_kernel_name(i).decode()] = (n.value, ms.value, items.value)
        return out

    # --- learned pair gate ---------------------------------------------------
    def mlp_set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32).ravel()
        assert p.size == lib().d2_mlp_num_params()
        self._check(lib().d2_mlp_set_params(self.h, p.ctypes.data_as(_fp)))

    def mlp_get_params(self):
        p = np.zeros(lib().d2_mlp_num_params(), dtype=np.float32)
        self._check(lib().d2_mlp_get_params(self.h, p.ctypes.data_as(_fp)))
        return p

    def mlp_infer(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, 24)
        Y = np.zeros((X.shape[0], 2), dtype=np.float32)
        self._check(lib().d2_mlp_infer(self.h, X.ctypes.data_as(_fp), X.shape[0], Y.ctypes.data_as(_fp)))
        return Y

    def mlp_loss_grad(self, X, T, lw=(1.0, 1.0), delta=0.1):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, 24)
        T = np.ascontiguousarray(T, dtype=np.float32).reshape(-1, 2)
        loss = np.zeros(1, dtype=np.float32)
        g = np.zeros(lib().d2_mlp_num_params(), dtype=np.float32)
        self._check(lib().d2_mlp_loss_grad(self.h, X.ctypes.data_as(_fp), T.ctypes.data_as(_fp), X.shape[0],
                                           float(lw[0]), float(lw[1]), float(delta), loss.ctypes.data_as(_fp),
                                           g.ctypes.data_as(_fp)))
        return float(loss[0]), g

    def mlp_train(self, X, T, epochs, batch=4096, lr=1e-3, lw=(1.0, 1.0), delta=0.1, seed=0):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, 24)
        T = np.ascontiguousarray(T, dtype=np.float32).reshape(-1, 2)
        out = np.zeros(epochs, dtype=np.float32)
        self._check(lib().d2_mlp_train(self.h, X.ctypes.data_as(_fp), T.ctypes.data_as(_fp), X.shape[0], int(epochs),
                                       int(batch), float(lr), float(lw[0]), float(lw[1]), float(delta), int(seed),
                                       out.ctypes.data_as(_fp)))
        return out

    def pair_features(self):
        X = np.zeros((self.n_pairs, 24), dtype=np.float32)
        self._check(lib().d2_pair_features(self.h, X.ctypes.data_as(_fp)))
        return X

    def set_learned(self, on=True, t_gate=1.1, s_gate=0.99, phases=3):
        """on: gate on/off; phases: mask of the phases it applies to (1 CCD, 2 discrete)."""
        self._check(lib().d2_set_learned(self.h, int(phases) if on else 0, float(t_gate), float(s_gate)))

    def gate_counts(self):
        """(proposed, certified) pairs of the CCD and of the discrete phase of the last run."""
        g, _ = self.run_info()
        return {"ccd": (int(g[4]), int(g[5])), "discrete": (int(g[6]), int(g[7]))}

    def run_info(self):
        """(global counters[8], per-level sizes[2, 17, 8]) of the last run:
        [0] the discrete phase (the CCD phase after ccd_run), [1] the last
        CCD phase."""
        buf = np.zeros(8 + 2 * 17 * 8, dtype=np.int64)
        lib().d2_run_info(self.h, buf.ctypes.data_as(_i64p), buf.size)
        return buf[:8], buf[8:].reshape(2, 17, 8)

    def sync(self):
        self._check(lib().d2_sync(self.h))

    def find_contacts_bucket(self, body_a, body_b, node_pairs, max_hits=None, entry="d2_find_contacts_bucket"):
        """findContactsBucketComparison over node pairs (node of a's tree,
        node of b's tree) at the bodies' future states: (pair_used[n],
        ints[h, 8], reals[h, 11]).  Raises CapacityError when more than
        max_hits hits exist (self.last_hits holds the count)."""
        nps = _i32(node_pairs).reshape(-1, 2)
        n = nps.shape[0]
        cap = max(64 * n, 1) if max_hits is None else int(max_hits)
        used = np.zeros(max(n, 1), dtype=np.int32)
        ints = np.zeros((max(cap, 1), 8), dtype=np.int32)
        reals = np.zeros((max(cap, 1), 11))
        hits = ctypes.c_int64(0)
        rc = getattr(lib(), entry)(self.h, int(body_a), int(body_b), n, _iptr(nps), _iptr(used), cap, _iptr(ints),
                                   _dptr(reals), ctypes.byref(hits))
        self.last_hits = hits.value
        self._check(rc)
        h = hits.value
        return used[:n], ints[:h], reals[:h]

    def filter_contacts(self, reals, margin=0.0):
        """filterContacts(unfiltered, margin) on the device: (sorted order of
        the inputs, kept inputs in sorted order)."""
        r = _f64(reals).reshape(-1, 11)
        n = r.shape[0]
        order = np.zeros(max(n, 1), dtype=np.int32)
        kept = np.zeros(max(n, 1), dtype=np.int32)
        k = ctypes.c_int(0)
        self._check(lib().d2_filter_contacts(self.h, n, _dptr(r), float(margin), _iptr(order), _iptr(kept),
                                             ctypes.byref(k)))
        return order[:n], kept[:k.value]

    def set_pruning(self, on=True):
        """Skip discrete node pairs whose leaves are provably apart (results unchanged)."""
        self._check(lib().d2_set_pruning(self.h, int(bool(on))))

    def set_touch_stats(self, on=True):
        """Accounting runs: record the (body, node) records each level tests."""
        self._check(lib().d2_set_touch_stats(self.h, int(bool(on))))

    def touch_stats(self):
        """(per_level[2, 17, 4], run[4]) of the last accounting run: distinct
        inner nodes, leaf nodes, leaf triangles and bodies, per phase (0
        discrete, 1 CCD) and level, and over the whole run."""
        n = lib().d2_touch_stats(self.h, None, 0)
        buf = np.zeros(n, dtype=np.int64)
        lib().d2_touch_stats(self.h, buf.ctypes.data_as(_i64p), n)
        return buf[:-4].reshape(2, 17, 4), buf[-4:]

    # --- broad phase and clusters ---------------------------------------------
    def upload_tree_bounds(self, aabb6, radius):
        box = np.ascontiguousarray(np.asarray(aabb6, dtype=np.float32).ravel())
        rad = _f64(radius)
        self._check(lib().d2_tree_bounds_upload(self.h, box.ctypes.data_as(_fp), _dptr(rad)))

    def tree_bounds(self, n_trees):
        box = np.zeros((n_trees, 6), dtype=np.float32)
        rad = np.zeros(n_trees)
        self._check(lib().d2_tree_bounds_download(self.h, box.ctypes.data_as(_fp), _dptr(rad)))
        return box, rad

    def upload_last_states(self, last14):
        st = _f64(last14)
        self._check(lib().d2_bodies_last_upload(self.h, _dptr(st)))

    def upload_groups(self, groups):
        g = None if groups is None else _i32(groups)
        self._check(lib().d2_bodies_groups_upload(self.h, None if g is None else _iptr(g)))

    def broad_phase(self, mode=1, t=0.0):
        """broadPhaseEmbree on the device (mode 0: with step t; mode 1: the
        individual-time form); the pairs become the resident pair list."""
        n = ctypes.c_int64(0)
        self._check(lib().d2_broad_phase(self.h, int(mode), float(t), ctypes.byref(n)))
        self.n_pairs = n.value
        return n.value

    def download_pairs(self):
        pairs = np.zeros((self.n_pairs, 2), dtype=np.int32)
        mt = np.zeros(self.n_pairs)
        self._check(lib().d2_pairs_download(self.h, _iptr(pairs), _dptr(mt)))
        return pairs, mt

    def build_clusters(self):
        """separateCollisionClusters on the device over the resident pairs."""
        k = ctypes.c_int(0)
        self._check(lib().d2_clusters_build(self.h, ctypes.byref(k)))
        self.n_clusters = k.value
        return k.value

    def download_clusters(self):
        k = self.n_clusters
        cop = np.zeros(max(self.n_pairs, 1), dtype=np.int32)
        cob = np.zeros(self.n_bodies, dtype=np.int32)
        first = np.zeros(max(k, 1), dtype=np.int32)
        mt = np.zeros(max(k, 1))
        st = np.zeros(max(k, 1))
        self._check(lib().d2_clusters_download(self.h, _iptr(cop), _iptr(cob), _iptr(first), _dptr(mt), _dptr(st)))
        return {"n": k, "of_pair": cop[:self.n_pairs], "of_body": cob, "first": first[:k], "min_time": mt[:k],
                "step": st[:k]}


# --- reference-named mirrors -------------------------------------------------
def compare_trees_full(ctx, pairs, filter=True):
    """compareTreesFull (+ filterContacts(.., 0.0)) for every pair, on the GPU."""
    ctx.upload_pairs(pairs)
    ctx.narrow_run(filter)
    return ctx.result()


def compare_trees_full_continuous(ctx, pairs, max_time):
    """compareTreesFullContinuous + TOI->step rule for every pair, on the GPU.

    Result.per_pair[:, 0] is the pair's step scale, [:, 1] its min_time_out."""
    ctx.upload_pairs(pairs, max_time)
    ctx.ccd_run()
    return ctx.result(with_pairs=True)


def fine_collision_clusters_with_time_step_selection(ctx, pairs, cluster_of_pair, n_clusters, last_step, dt):
    """Per-cluster step after the CCD (separate_clusters.cpp:462-599 + clamp)."""
    pairs = _i32(pairs).reshape(-1, 2)
    last_step = np.asarray(last_step, dtype=np.float64)
    mt = np.minimum(last_step[pairs[:, 0]], last_step[pairs[:, 1]])
    ctx.upload_pairs(pairs, mt)
    ctx.ccd_run()
    return ctx.cluster_steps(cluster_of_pair, n_clusters, dt)
