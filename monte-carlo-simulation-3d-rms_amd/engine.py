"""Python front end of the HIP engine (librms_amd.so).

`Engine` owns one rms_ctx (one HIP stream on one device); `Engine.plan`
validates a list of jobs -- one job is one reference setup directory: a
geometry, its walkers and its b-table or PGSE table -- and uploads their
tables once; `Plan.execute` runs all jobs in one persistent launch and returns
the record sums of hpc_code/lib/main.cu:341-378 (and main_PGSE.cu:348-357),
split into the reference's output arrays.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi as capi


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Geometry:
    """Voxelised label lattice plus per-class physics.

    labels: (nx, ny, nz) uint8/uint16, x-major with z fastest (the order of
    fiber.txt, analysis/setupRMS.m:17-23).  Class 0 is the null compartment.
    D: per-class diffusivity (um^2/ms); T2: per-class T2 (ms, inf = none);
    kappa: permeability between non-null classes (um/ms): None keeps the
    reference rule P = min(1, sqrt(Dj/Di)) (main.cu:444-458), a scalar applies
    to every pair, a matrix per pair.  label_class optionally maps labels to
    classes (uint16 lattices with many labels).  periodic names the axes
    ("x", "y", "z" in any combination) whose lattice edges wrap around; the
    others are the reference's mirrors (main.cu:246-265).  storage: "auto",
    "dense" or "bricks", the device layout of the cell words
    (rms_geometry.storage; results are identical)."""

    STORAGE = {"auto": 0, "dense": 1, "bricks": 2}

    def __init__(self, labels, res, D, T2=None, kappa=None, label_class=None, periodic="", storage="auto"):
        labels = np.asarray(labels)
        if labels.ndim != 3:
            raise ValueError("labels must be a 3-D array")
        if labels.dtype not in (np.uint8, np.uint16):
            if labels.min() < 0 or labels.max() > 65535:
                raise ValueError("labels must be in 0..65535")
            labels = labels.astype(np.uint16 if labels.max() > 255 else np.uint8)
        self.labels = np.ascontiguousarray(labels)
        self.res = float(res)
        self.D = [float(x) for x in D]
        nc = len(self.D)
        if not 1 <= nc <= capi.NC_MAX:
            raise ValueError(f"1..{capi.NC_MAX} classes supported")
        self.T2 = [float(x) for x in (T2 if T2 is not None else [np.inf] * nc)]
        K = np.full((capi.NC_MAX, capi.NC_MAX), -1.0)
        if kappa is not None:
            kk = np.asarray(kappa, dtype=np.float64)
            if kk.ndim == 0:
                for i in range(1, nc):
                    for j in range(1, nc):
                        if i != j:
                            K[i, j] = float(kk)
            else:
                K[: kk.shape[0], : kk.shape[1]] = kk
        self.kappa = K
        self.label_class = None if label_class is None else np.ascontiguousarray(label_class, np.uint8)
        if any(a not in "xyz" for a in periodic):
            raise ValueError("periodic: axes among 'x', 'y', 'z'")
        self.periodic = sum(1 << "xyz".index(a) for a in set(periodic))
        if storage not in self.STORAGE:
            raise ValueError("storage: 'auto', 'dense' or 'bricks'")
        self.storage = storage

    @property
    def nclass(self):
        return len(self.D)

    @property
    def shape(self):
        return self.labels.shape

    def to_c(self):
        g = capi.Geometry()
        g.nx, g.ny, g.nz = self.labels.shape
        g.label_bytes = self.labels.dtype.itemsize
        g.labels = _ptr(self.labels)
        g.label_class = _ptr(self.label_class)
        g.nlabel = 0 if self.label_class is None else len(self.label_class)
        g.nclass = self.nclass
        g.res = self.res
        for i in range(self.nclass):
            g.D[i] = self.D[i]
            g.T2[i] = self.T2[i]
        for i in range(capi.NC_MAX * capi.NC_MAX):
            g.kappa[i] = float(self.kappa.flat[i])
        g.periodic = self.periodic
        g.storage = self.STORAGE[self.storage]
        return g

    def derive(self, dt):
        """step (voxels), T2 (steps) and P per class pair, main.cu:429-458."""
        st = np.zeros(capi.NC_MAX)
        t2 = np.zeros(capi.NC_MAX)
        P = np.zeros(capi.NC_MAX * capi.NC_MAX)
        g = self.to_c()
        s = capi.lib().rms_derive(C.byref(g), float(dt), _ptr(st), _ptr(t2), _ptr(P))
        if s != 0:
            raise capi.RmsError(s, "rms_derive")
        n = self.nclass
        return {"step": st[:n], "T2s": t2[:n], "P": P.reshape(capi.NC_MAX, capi.NC_MAX)[:n, :n]}


def nval(nclass, nbvec):
    return 1 + nclass + 6 + 15 + nbvec


def split_records(r, nclass, nbvec):
    """Record matrix -> the reference's sig0 / NPar_count / dx2 / dx4 / sigRe."""
    rec = r["rec"]
    r["sig0"] = rec[:, 0]
    r["NPar_count"] = rec[:, 1:1 + nclass]
    r["dx2"] = rec[:, 1 + nclass:7 + nclass]
    r["dx4"] = rec[:, 7 + nclass:22 + nclass]
    r["sigRe"] = rec[:, 22 + nclass:22 + nclass + nbvec]
    return r


class Engine:
    """One device context (rms_create)."""

    def __init__(self, device=0):
        L = capi.lib()
        h = C.c_void_p()
        opts = capi.DeviceOpts(device=int(device), reserved=0)
        s = L.rms_create(C.byref(opts), C.byref(h))
        if s != 0:
            raise capi.RmsError(s, "rms_create failed (no HIP device?)")
        self._h = h
        self._geom_ids = {}

    def _check(self, s, what):
        if s != 0:
            msg = capi.lib().rms_last_error(self._h).decode()
            raise capi.RmsError(s, f"{what}: {msg}")

    def close(self):
        if self._h:
            capi.lib().rms_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_geometry(self, geom: Geometry) -> int:
        key = id(geom)
        if key in self._geom_ids and self._geom_ids[key][1] is geom:
            return self._geom_ids[key][0]
        g = geom.to_c()
        gid = C.c_int32()
        self._check(capi.lib().rms_add_geometry(self._h, C.byref(g), C.byref(gid)), "rms_add_geometry")
        self._geom_ids[key] = (gid.value, geom)
        return gid.value

    def geometry_info(self, geom: Geometry, wrap_cells=False):
        """(device bytes, "dense" or "bricks") of an uploaded geometry, and
        with wrap_cells whether steps through its periodic sides take the
        fast path (the rule of rms_geometry_info)."""
        gid = self.add_geometry(geom)
        nb, st, wc = C.c_int64(), C.c_int32(), C.c_int32()
        self._check(capi.lib().rms_geometry_info(self._h, gid, C.byref(nb), C.byref(st), C.byref(wc)),
                    "rms_geometry_info")
        out = (int(nb.value), {1: "dense", 2: "bricks"}[st.value])
        return out + (bool(wc.value),) if wrap_cells else out

    def plan(self, jobs, dt, TN, timepoints=1000, precision="f32", replicas=32, chunk_steps=0, use_graph=False,
             block_threads=0, grid_blocks=0, xcd_split=0, counters=True):
        """counters=False leaves the diagnostic counters (crossings,
        reflections, mirrors, draws) at zero: about 3 % less work per step."""
        return Plan(self, jobs, dt, TN, timepoints, precision, replicas, chunk_steps, use_graph, block_threads,
                    grid_blocks, xcd_split, counters)

    def run_noalloc(self, job, dt, TN, timepoints=1000, precision="f32", replicas=32, block_threads=0, grid_blocks=0,
                    xcd_split=0, counters=True):
        """One job through rms_run_noalloc: the walk of `run` with every device
        buffer in the library's static scratch array (no device allocation;
        the geometry is uploaded once, on first use, like `run` does)."""
        geom = job["geom"]
        gid = self.add_geometry(geom)
        bval, bvec, grad = job.get("bval"), job.get("bvec"), job.get("grad")
        bval = None if bval is None else np.ascontiguousarray(bval, np.float64).ravel()
        bvec = None if bvec is None else np.ascontiguousarray(bvec, np.float64).reshape(-1, 3)
        grad = None if grad is None else np.ascontiguousarray(grad, np.float64).reshape(-1, 6)
        nb = 0 if bval is None else len(bval)
        ng = 0 if grad is None else grad.shape[0]
        cj = capi.Job(geom_id=gid, nwalk=int(job["nwalk"]), seed=int(job.get("seed", 1)) & 0xFFFFFFFFFFFFFFFF,
                      nbvec=nb, bval=_ptr(bval), bvec=_ptr(bvec), ngrad=ng, grad=_ptr(grad),
                      walker_base=int(job.get("walker_base", 0)))
        R = max(1, int(replicas))
        prm = capi.RunParams(dt=float(dt), TN=int(TN), timepoints=int(timepoints), precision=capi.PREC[precision],
                             replicas=R, chunk_steps=0, use_graph=0, block_threads=int(block_threads),
                             grid_blocks=int(grid_blocks), xcd_split=int(xcd_split), skip_counters=int(not counters))
        nv = nval(geom.nclass, nb)
        r = {"rec": np.zeros((int(timepoints), nv)), "rec_rep": np.zeros((R, int(timepoints), nv)),
             "sig": np.zeros(ng) if ng else None, "sig_rep": np.zeros((R, ng)) if ng else None}
        co = capi.JobOut()
        co.rec, co.rec_rep = _ptr(r["rec"]), _ptr(r["rec_rep"])
        co.sig, co.sig_rep = _ptr(r["sig"]), _ptr(r["sig_rep"])
        tm = capi.Timing()
        self._check(capi.lib().rms_run_noalloc(self._h, C.byref(cj), C.byref(prm), C.byref(co), C.byref(tm)),
                    "rms_run_noalloc")
        r["counters"] = np.array(co.counters[:], np.uint64)
        r["timing"] = {f: getattr(tm, f) for f, _ in capi.Timing._fields_}
        return split_records(r, geom.nclass, nb)

    def run(self, jobs, dt, TN, timepoints=1000, precision="f32", replicas=32, final_state=False, **kw):
        p = self.plan(jobs, dt, TN, timepoints, precision, replicas, **kw)
        try:
            return p.execute(final_state=final_state)
        finally:
            p.close()


class Plan:
    """rms_plan_create / rms_plan_execute / rms_plan_adjoint."""

    def __init__(self, eng, jobs, dt, TN, timepoints, precision, replicas, chunk_steps, use_graph, block_threads,
                 grid_blocks, xcd_split, counters=True):
        self.eng = eng
        self.jobs = []
        n = len(jobs)
        cj = (capi.Job * n)()
        keep = []
        for k, jb in enumerate(jobs):
            geom = jb["geom"]
            gid = eng.add_geometry(geom)
            bval = jb.get("bval")
            bvec = jb.get("bvec")
            grad = jb.get("grad")
            bval = None if bval is None else np.ascontiguousarray(bval, np.float64).ravel()
            bvec = None if bvec is None else np.ascontiguousarray(bvec, np.float64).reshape(-1, 3)
            grad = None if grad is None else np.ascontiguousarray(grad, np.float64).reshape(-1, 6)
            nb = 0 if bval is None else len(bval)
            if nb and bvec.shape[0] != nb:
                raise ValueError("bval and bvec lengths differ")
            ng = 0 if grad is None else grad.shape[0]
            keep += [bval, bvec, grad]
            cj[k].geom_id = gid
            cj[k].nwalk = int(jb["nwalk"])
            cj[k].seed = int(jb.get("seed", 1)) & 0xFFFFFFFFFFFFFFFF
            cj[k].nbvec = nb
            cj[k].bval = _ptr(bval)
            cj[k].bvec = _ptr(bvec)
            cj[k].ngrad = ng
            cj[k].grad = _ptr(grad)
            cj[k].walker_base = int(jb.get("walker_base", 0))
            self.jobs.append(dict(nwalk=int(jb["nwalk"]), nclass=geom.nclass, nbvec=nb, ngrad=ng))
        self.any_nb = any(j["nbvec"] for j in self.jobs)
        self.params = capi.RunParams(
            dt=float(dt), TN=int(TN), timepoints=int(timepoints), precision=capi.PREC[precision],
            replicas=max(1, int(replicas)), chunk_steps=int(chunk_steps), use_graph=int(bool(use_graph)),
            block_threads=int(block_threads), grid_blocks=int(grid_blocks), xcd_split=int(xcd_split),
            skip_counters=int(not counters))
        h = C.c_void_p()
        eng._check(capi.lib().rms_plan_create(eng._h, cj, n, C.byref(self.params), C.byref(h)), "rms_plan_create")
        self._h = h
        self.timing = None

    def close(self):
        if getattr(self, "_h", None):
            capi.lib().rms_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def execute(self, final_state=False, want_replicas=True):
        P = self.params
        T = P.timepoints
        R = P.replicas
        n = len(self.jobs)
        co = (capi.JobOut * n)()
        res = []
        for k, j in enumerate(self.jobs):
            nb = j["nbvec"] if self.any_nb else 0
            nv = nval(j["nclass"], j["nbvec"])
            ng = j["ngrad"]
            r = {"rec": np.zeros((T, nv)), "sig": np.zeros(ng) if ng else None}
            if want_replicas:
                r["rec_rep"] = np.zeros((R, T, nv))
                r["sig_rep"] = np.zeros((R, ng)) if ng else None
            co[k].rec = _ptr(r["rec"])
            co[k].sig = _ptr(r["sig"])
            if want_replicas:
                co[k].rec_rep = _ptr(r["rec_rep"])
                co[k].sig_rep = _ptr(r["sig_rep"])
            if final_state:
                W = j["nwalk"]
                r["fin_g"] = np.zeros((3, W), np.int32)
                r["fin_o"] = np.zeros((3, W))
                r["fin_gi"] = np.zeros((3, W), np.int32)
                r["fin_oi"] = np.zeros((3, W))
                r["fin_c"] = np.zeros(W, np.int32)
                r["fin_cnt"] = np.zeros((5, W), np.uint32)
                for name in ("fin_g", "fin_o", "fin_gi", "fin_oi", "fin_c", "fin_cnt"):
                    setattr(co[k], name, _ptr(r[name]))
            del nb
            res.append(r)
        tm = capi.Timing()
        self.eng._check(capi.lib().rms_plan_execute(self._h, co, C.byref(tm)), "rms_plan_execute")
        self.timing = {f: getattr(tm, f) for f, _ in capi.Timing._fields_}
        for k, (r, j) in enumerate(zip(res, self.jobs)):
            r["counters"] = np.array(co[k].counters[:], np.uint64)
            split_records(r, j["nclass"], j["nbvec"])
        return res

    def adjoint(self, cot_rec, cot_sig=None, want_replicas=False, estimator="auto", fd_step=0.0,
                fd_all_walkers=False):
        """Vector-Jacobian product of the record/signal sums with respect to
        D[c] (entries 0..3), T2[c] (4..7) and kappa (8).  estimator picks the
        D entries (rms_adjoint_opts): "auto" pathwise where the walk is
        continuous in D, central differences with common random numbers
        elsewhere; "fd" central differences for every job; "boundary"
        pathwise plus the boundary terms; "pathwise" pathwise only (NaN where
        the walk jumps).  fd_step is the relative step h (0 = 0.05).  The
        central differences for class c walk only the walkers whose walk can
        change with D_c, unless fd_all_walkers."""
        n = len(self.jobs)
        ca = (capi.JobAdj * n)()
        outs = []
        keep = []
        for k, j in enumerate(self.jobs):
            cr = None if cot_rec is None or cot_rec[k] is None else np.ascontiguousarray(cot_rec[k], np.float64)
            cs = None if cot_sig is None or cot_sig[k] is None else np.ascontiguousarray(cot_sig[k], np.float64)
            keep += [cr, cs]
            g = np.zeros(capi.NPARAM)
            gr = np.zeros((self.params.replicas, capi.NPARAM)) if want_replicas else None
            ca[k].cot_rec = _ptr(cr)
            ca[k].cot_sig = _ptr(cs)
            ca[k].grad = _ptr(g)
            ca[k].grad_rep = _ptr(gr)
            outs.append({"grad": g, "grad_rep": gr})
        tm = capi.Timing()
        opts = capi.AdjointOpts(d_estimator=capi.D_ESTIMATOR[estimator], fd_all_walkers=int(bool(fd_all_walkers)),
                                fd_rel_step=float(fd_step))
        self.eng._check(capi.lib().rms_plan_adjoint_ex(self._h, ca, C.byref(opts), C.byref(tm)),
                        "rms_plan_adjoint_ex")
        self.adj_timing = {f: getattr(tm, f) for f, _ in capi.Timing._fields_}
        return outs
