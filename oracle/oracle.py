"""ctypes front end of the CPU oracle (liboracle_rms.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.  It restates the
reference engine (hpc_code/lib/main.cu:63-384, main_PGSE.cu:336-357); the C
sources carry the line-by-line citations.
"""
from __future__ import annotations

import copy
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_rms.so")
NC_MAX = 4
BATCH = 64
PREC = {"f32": 0, "f64": 1}


class _Geom(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("label_bytes", C.c_int32),
        ("labels", C.c_void_p),
        ("label_class", C.c_void_p),
        ("nlabel", C.c_int32),
        ("nclass", C.c_int32),
        ("res", C.c_double),
        ("D", C.c_double * NC_MAX),
        ("T2", C.c_double * NC_MAX),
        ("kappa", C.c_double * (NC_MAX * NC_MAX)),
        ("periodic", C.c_int32),
        ("wrap_cells", C.c_int32),
    ]


class _Derived(C.Structure):
    _fields_ = [
        ("nclass", C.c_int32),
        ("step", C.c_double * NC_MAX),
        ("inv_step", C.c_double * NC_MAX),
        ("T2s", C.c_double * NC_MAX),
        ("P", C.c_double * (NC_MAX * NC_MAX)),
        ("Pcat", C.c_int32 * (NC_MAX * NC_MAX)),
    ]


class _Job(C.Structure):
    _fields_ = [
        ("geom", C.POINTER(_Geom)),
        ("nwalk", C.c_int64),
        ("seed", C.c_uint64),
        ("nbvec", C.c_int32),
        ("bval", C.c_void_p),
        ("bvec", C.c_void_p),
        ("ngrad", C.c_int32),
        ("grad", C.c_void_p),
        ("walker_base", C.c_int64),
        ("wlist", C.c_void_p),
    ]


class _Params(C.Structure):
    _fields_ = [
        ("dt", C.c_double),
        ("TN", C.c_int64),
        ("timepoints", C.c_int32),
        ("prec", C.c_int32),
        ("replicas", C.c_int32),
        ("pgse_bruteforce", C.c_int32),
        ("nthreads", C.c_int32),
        ("no_boundary", C.c_int32),
        ("reserved", C.c_int32),
    ]


class _Out(C.Structure):
    _fields_ = [
        ("rec", C.c_void_p),
        ("rec_rep", C.c_void_p),
        ("sig", C.c_void_p),
        ("sig_rep", C.c_void_p),
        ("counters", C.c_uint64 * 5),
        ("fin_g", C.c_void_p),
        ("fin_o", C.c_void_p),
        ("fin_gi", C.c_void_p),
        ("fin_oi", C.c_void_p),
        ("fin_c", C.c_void_p),
        ("fin_cnt", C.c_void_p),
        ("fin_vis", C.c_void_p),
    ]


class _Adj(C.Structure):
    _fields_ = [
        ("cot_rec", C.c_void_p),
        ("cot_sig", C.c_void_p),
        ("grad", C.c_void_p),
        ("grad_rep", C.c_void_p),
    ]


_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc only)."""
    args = ["make", "-s", "-C", HERE]
    if force:
        args.insert(1, "-B")
    subprocess.run(args, check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.or_philox4x32_10.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_u01_f64.argtypes = [C.c_uint32, C.c_uint32]
        L.or_u01_f64.restype = C.c_double
        L.or_u01_f32.argtypes = [C.c_uint32]
        L.or_u01_f32.restype = C.c_float
        L.or_sincos_turns_f64.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.or_sincos_turns_f32.argtypes = [C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.or_exp_f64.argtypes = [C.c_double]
        L.or_exp_f64.restype = C.c_double
        L.or_sincos_f64.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.or_cos_f64.argtypes = [C.c_double]
        L.or_cos_f64.restype = C.c_double
        L.or_wrap_cells_rule.argtypes = [C.POINTER(_Geom)]
        L.or_wrap_cells_rule.restype = C.c_int32
        L.or_derive.argtypes = [C.POINTER(_Geom), C.c_double, C.POINTER(_Derived)]
        L.or_derive.restype = C.c_int
        L.or_nval.argtypes = [C.c_int32, C.c_int32]
        L.or_nval.restype = C.c_int32
        L.or_diff_times.argtypes = [C.c_double, C.c_int64, C.c_int32, C.c_void_p]
        L.or_pgse_lobes.argtypes = [C.c_double, C.c_int64, C.c_int32, C.c_void_p, C.c_void_p]
        L.or_march_f64.argtypes = [C.POINTER(_Geom), C.POINTER(_Derived)] + [C.c_void_p] * 8
        L.or_march_f64.restype = C.c_int
        L.or_fast_step_f64.argtypes = [C.POINTER(_Geom), C.POINTER(_Derived)] + [C.c_void_p] * 3 + \
            [C.c_int32] + [C.c_void_p] * 4
        L.or_fast_step_f64.restype = C.c_int
        L.or_run.argtypes = [C.POINTER(_Job), C.c_int32, C.POINTER(_Params), C.POINTER(_Out)]
        L.or_run.restype = C.c_int
        L.or_run_adjoint.argtypes = [C.POINTER(_Job), C.c_int32, C.POINTER(_Params), C.POINTER(_Adj),
                                     C.POINTER(_Out)]
        L.or_run_adjoint.restype = C.c_int
        L.or_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


# ---------------------------------------------------------------- primitives
def philox(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, np.uint32)
    lib().or_philox4x32_10(_ptr(c), _ptr(k), _ptr(out))
    return out


def u01_f64(a, b):
    return lib().or_u01_f64(int(a), int(b))


def u01_f32(w):
    return lib().or_u01_f32(int(w))


def sincos_turns(t, prec="f64"):
    if prec == "f64":
        s, c = C.c_double(), C.c_double()
        lib().or_sincos_turns_f64(float(t), C.byref(s), C.byref(c))
    else:
        s, c = C.c_float(), C.c_float()
        lib().or_sincos_turns_f32(float(t), C.byref(s), C.byref(c))
    return s.value, c.value


def exp(x):
    return lib().or_exp_f64(float(x))


def sincos(x):
    s, c = C.c_double(), C.c_double()
    lib().or_sincos_f64(float(x), C.byref(s), C.byref(c))
    return s.value, c.value


# ---------------------------------------------------------------- geometry
class Geometry:
    """Label lattice plus per-class physics (the fiber/diffusivity/T2 inputs)."""

    def __init__(self, labels, res, D, T2=None, kappa=None, label_class=None, periodic=""):
        labels = np.ascontiguousarray(labels)
        if labels.dtype not in (np.uint8, np.uint16):
            labels = labels.astype(np.uint16 if labels.max() > 255 else np.uint8)
        self.labels = labels
        self.res = float(res)
        self.D = [float(x) for x in D]
        nc = len(self.D)
        self.T2 = [float(x) for x in (T2 if T2 is not None else [np.inf] * nc)]
        K = np.full((NC_MAX, NC_MAX), -1.0)
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
        self.periodic = sum(1 << "xyz".index(a) for a in set(periodic))

    @property
    def nclass(self):
        return len(self.D)

    def _c(self):
        g = _Geom()
        g.nx, g.ny, g.nz = self.labels.shape
        g.label_bytes = self.labels.dtype.itemsize
        g.labels = _ptr(self.labels)
        g.label_class = _ptr(self.label_class) if self.label_class is not None else None
        g.nlabel = 0 if self.label_class is None else len(self.label_class)
        g.nclass = self.nclass
        g.res = self.res
        for i in range(self.nclass):
            g.D[i] = self.D[i]
            g.T2[i] = self.T2[i]
        for i in range(NC_MAX * NC_MAX):
            g.kappa[i] = self.kappa.flat[i]
        g.periodic = self.periodic
        g.wrap_cells = lib().or_wrap_cells_rule(C.byref(g))
        return g


def derive(geom: Geometry, dt: float):
    g = geom._c()
    d = _Derived()
    if not lib().or_derive(C.byref(g), dt, C.byref(d)):
        raise ValueError(lib().or_last_error().decode())
    nc = geom.nclass
    P = np.array(d.P[:]).reshape(NC_MAX, NC_MAX)[:nc, :nc]
    return {
        "step": np.array(d.step[:nc]),
        "T2s": np.array(d.T2s[:nc]),
        "P": P,
        "Pcat": np.array(d.Pcat[:]).reshape(NC_MAX, NC_MAX)[:nc, :nc],
        "_c": d,
    }


def diff_times(dt, TN, timepoints):
    TD = np.zeros(timepoints)
    lib().or_diff_times(dt, TN, timepoints, _ptr(TD))
    return TD


def pgse_lobes(dt, TN, grad):
    grad = np.ascontiguousarray(grad, np.float64).reshape(-1, 6)
    abe = np.zeros((grad.shape[0], 3), np.int64)
    lib().or_pgse_lobes(dt, TN, grad.shape[0], _ptr(grad), _ptr(abe))
    return abe


def march_f64(geom: Geometry, dt, g, o, v, cls, u_exch=(0.5, 0.5, 0.5, 0.5), gi=None, oi=None):
    """One face march (main.cu:153-300) from voxel g, offset o, direction v."""
    gc = geom._c()
    d = _Derived()
    if not lib().or_derive(C.byref(gc), dt, C.byref(d)):
        raise ValueError(lib().or_last_error().decode())
    g = np.array(g, np.int32)
    o = np.array(o, np.float64)
    v = np.array(v, np.float64)
    c = np.array([cls], np.int32)
    u = np.array(u_exch, np.float64)
    cnt = np.zeros(4, np.int32)
    gi = np.array(g if gi is None else gi, np.int32)
    oi = np.array(o if oi is None else oi, np.float64)
    nd = lib().or_march_f64(C.byref(gc), C.byref(d), _ptr(g), _ptr(o), _ptr(v), _ptr(c), _ptr(u), _ptr(cnt),
                            _ptr(gi), _ptr(oi))
    return dict(g=g, o=o, v=v, cls=int(c[0]), cnt=cnt, draws=nd, gi=gi, oi=oi)


def fast_step_f64(geom: Geometry, dt, g, o, v, cls, u_exch=(0.5, 0.5, 0.5, 0.5), gi=None, oi=None):
    """One straight-move fast-path step; None when the march is needed.
    cnt = crossings, reflections, mirrors, exchange draws (draw j is
    u_exch[j]); gi, oi = the (mirror-unfolded) origin, default the start."""
    gc = geom._c()
    d = _Derived()
    if not lib().or_derive(C.byref(gc), dt, C.byref(d)):
        raise ValueError(lib().or_last_error().decode())
    g = np.array(g, np.int32)
    o = np.array(o, np.float64)
    v = np.array(v, np.float64)
    u = np.array(u_exch, np.float64)
    cnt = np.zeros(4, np.int32)
    gi = np.array(g if gi is None else gi, np.int32)
    oi = np.array(o if oi is None else oi, np.float64)
    ok = lib().or_fast_step_f64(C.byref(gc), C.byref(d), _ptr(g), _ptr(o), _ptr(v), int(cls), _ptr(u), _ptr(cnt),
                                _ptr(gi), _ptr(oi))
    return dict(g=g, o=o, cnt=cnt, gi=gi, oi=oi) if ok else None


# ---------------------------------------------------------------- simulation
def nval(nclass, nbvec):
    return 1 + nclass + 6 + 15 + nbvec


def run(jobs, dt, TN, timepoints, prec="f32", replicas=32, pgse_bruteforce=False, nthreads=0,
        final_state=False, cot_rec=None, cot_sig=None, boundary=True, marks=False):
    """Run a list of jobs; each job is a dict(geom, nwalk, seed, bval, bvec, grad,
    optionally walker_base and wlist: walk walkers wlist[k] of the job).

    Returns per job a dict with the record totals split into the
    reference's output arrays (sig0, NPar_count, dx2, dx4, sigRe), the PGSE
    signal, replica sums, counters and optionally the final walker state.
    With cotangents, the adjoint pass; boundary=False leaves out the
    boundary terms of the D derivative (the tangent pass alone); marks=True
    returns its per-walker marks "vis" (bit c = visited class c, bit
    4 + 4a + b = drew at an exchange face from class a into class b)."""
    L = lib()
    n = len(jobs)
    cj = (_Job * n)()
    co = (_Out * n)()
    keep = []
    res = []
    for k, jb in enumerate(jobs):
        geom = jb["geom"]
        gc = geom._c()
        keep.append(gc)
        bval = jb.get("bval")
        bvec = jb.get("bvec")
        grad = jb.get("grad")
        nb = 0 if bval is None else len(bval)
        bval = None if bval is None else np.ascontiguousarray(bval, np.float64)
        bvec = None if bvec is None else np.ascontiguousarray(bvec, np.float64).reshape(-1, 3)
        ng = 0 if grad is None else np.asarray(grad).reshape(-1, 6).shape[0]
        grad = None if grad is None else np.ascontiguousarray(grad, np.float64).reshape(-1, 6)
        keep += [bval, bvec, grad]
        cj[k].geom = C.pointer(gc)
        cj[k].nwalk = int(jb["nwalk"])
        cj[k].seed = int(jb.get("seed", 1)) & 0xFFFFFFFFFFFFFFFF
        cj[k].nbvec = nb
        cj[k].bval = _ptr(bval)
        cj[k].bvec = _ptr(bvec)
        cj[k].ngrad = ng
        cj[k].grad = _ptr(grad)
        cj[k].walker_base = int(jb.get("walker_base", 0))
        wl = jb.get("wlist")
        if wl is not None:
            wl = np.ascontiguousarray(wl, np.uint32)
            if wl.size != int(jb["nwalk"]):
                raise ValueError("wlist needs nwalk entries")
            keep.append(wl)
        cj[k].wlist = _ptr(wl)
        nv = nval(geom.nclass, nb)
        R = max(1, replicas)
        r = {
            "rec": np.zeros((timepoints, nv)),
            "rec_rep": np.zeros((R, timepoints, nv)),
            "sig": np.zeros(ng) if ng else None,
            "sig_rep": np.zeros((R, ng)) if ng else None,
        }
        co[k].rec = _ptr(r["rec"])
        co[k].rec_rep = _ptr(r["rec_rep"])
        co[k].sig = _ptr(r["sig"]) if ng else None
        co[k].sig_rep = _ptr(r["sig_rep"]) if ng else None
        if final_state:
            W = int(jb["nwalk"])
            r["fin_g"] = np.zeros((3, W), np.int32)
            r["fin_o"] = np.zeros((3, W))
            r["fin_gi"] = np.zeros((3, W), np.int32)
            r["fin_oi"] = np.zeros((3, W))
            r["fin_c"] = np.zeros(W, np.int32)
            r["fin_cnt"] = np.zeros((5, W), np.uint32)
            for name in ("fin_g", "fin_o", "fin_gi", "fin_oi", "fin_c", "fin_cnt"):
                setattr(co[k], name, _ptr(r[name]))
        if marks:
            r["vis"] = np.zeros(int(jb["nwalk"]), np.uint32)
            co[k].fin_vis = _ptr(r["vis"])
        r["_nclass"] = geom.nclass
        r["_nbvec"] = nb
        res.append(r)
    p = _Params(dt=float(dt), TN=int(TN), timepoints=int(timepoints), prec=PREC[prec], replicas=max(1, replicas),
                pgse_bruteforce=int(bool(pgse_bruteforce)), nthreads=int(nthreads), no_boundary=int(not boundary))
    if cot_rec is None and cot_sig is None:
        if not L.or_run(cj, n, C.byref(p), co):
            raise RuntimeError("oracle: " + L.or_last_error().decode())
    else:
        ca = (_Adj * n)()
        for k in range(n):
            cr = None if cot_rec is None or cot_rec[k] is None else np.ascontiguousarray(cot_rec[k], np.float64)
            cs = None if cot_sig is None or cot_sig[k] is None else np.ascontiguousarray(cot_sig[k], np.float64)
            g = np.zeros(16)
            gr = np.zeros((max(1, replicas), 16))
            keep += [cr, cs]
            ca[k].cot_rec = _ptr(cr)
            ca[k].cot_sig = _ptr(cs)
            ca[k].grad = _ptr(g)
            ca[k].grad_rep = _ptr(gr)
            res[k]["grad"] = g
            res[k]["grad_rep"] = gr
        if not L.or_run_adjoint(cj, n, C.byref(p), ca, co):
            raise RuntimeError("oracle: " + L.or_last_error().decode())
    for k, r in enumerate(res):
        r["counters"] = np.array(co[k].counters[:], np.uint64)
        split_records(r, r["_nclass"], r["_nbvec"])
    return res


def split_records(r, nclass, nbvec):
    """Split the record matrix into the reference's output arrays."""
    rec = r["rec"]
    r["sig0"] = rec[:, 0]
    r["NPar_count"] = rec[:, 1:1 + nclass]
    r["dx2"] = rec[:, 1 + nclass:7 + nclass]
    r["dx4"] = rec[:, 7 + nclass:22 + nclass]
    r["sigRe"] = rec[:, 22 + nclass:22 + nclass + nbvec]
    return r


# ------------------------------------------------- central differences of D
def fd_walker_mask(geom: Geometry, dt: float, c: int, h: float):
    """Marks (bits of the adjoint pass's per-walker "vis") that make a walker's
    walk depend on D_c between D_c (1 - h) and D_c (1 + h): class c, and the
    draws at exchange faces whose P differs at one of the three points.  None
    when such a face crosses or reflects without a draw at one of them (its
    walkers carry no mark): every walker is needed."""
    def P_at(f):
        g = copy.copy(geom)
        g.D = list(geom.D)
        g.D[c] = geom.D[c] * f
        d = derive(g, dt)
        return d["P"], d["Pcat"]

    P0, K0 = P_at(1.0)
    Pp, Kp = P_at(1.0 + h)
    Pm, Km = P_at(1.0 - h)
    mask = 1 << c
    for a in range(1, geom.nclass):
        for b in range(1, geom.nclass):
            if a == b or (Pp[a, b] == P0[a, b] and Pm[a, b] == P0[a, b]):
                continue
            if K0[a, b] != 2 or Kp[a, b] != 2 or Km[a, b] != 2:
                return None
            mask |= 1 << (4 + a * NC_MAX + b)
    return mask


def contract(cot_rec, rec, cot_sig=None, sig=None):
    """sum(cot_rec * rec) + sum(cot_sig * sig), added in row-major order."""
    f = 0.0
    for x, y in zip(np.ravel(cot_rec).tolist(), np.ravel(rec).tolist()):
        f = f + x * y
    if cot_sig is not None and sig is not None:
        for x, y in zip(np.ravel(cot_sig).tolist(), np.ravel(sig).tolist()):
            f = f + x * y
    return f


def central_differences(jobs, dt, TN, timepoints, cot_rec, cot_sig=None, h=0.05, prec="f32", replicas=32,
                        nthreads=0, subsets=True):
    """D entries of the adjoint by central differences with common random
    numbers: per non-null class c, forward runs with D_c (1 +- h) and the
    same seeds, dF/dD_c = (F+ - F-) / (D+ - D-) per job, total and per
    replica.  With subsets the runs for class c walk, per job, only the
    walkers whose adjoint-pass marks meet fd_walker_mask (in index order;
    none: the entry is 0), from one tangent pass (boundary=False) run here.
    Returns (grad[njobs][4], grad_rep[njobs][replicas][4]); entry 0 unused."""
    n = len(jobs)
    R = max(1, replicas)
    grad = np.zeros((n, 4))
    grad_rep = np.zeros((n, R, 4))
    marks = None
    if subsets:
        marks = [r["vis"] for r in run(jobs, dt, TN, timepoints, prec=prec, replicas=R, nthreads=nthreads,
                                       cot_rec=cot_rec, cot_sig=cot_sig, boundary=False, marks=True)]
    maxc = max(j["geom"].nclass for j in jobs)
    for c in range(1, maxc):
        todo = []
        for k, jb in enumerate(jobs):
            if jb["geom"].nclass <= c:
                continue
            sel = dict(jb)
            if subsets:
                mask = fd_walker_mask(jb["geom"], dt, c, h)
                if mask is not None:
                    idx = np.nonzero(marks[k] & np.uint32(mask))[0].astype(np.uint32)
                    if idx.size == 0:
                        continue  # entry 0: no walker can walk differently
                    sel.update(nwalk=int(idx.size), wlist=idx)
            todo.append((k, sel))
        if not todo:
            continue
        F, Dv = [], []
        for f in (1.0 + h, 1.0 - h):
            pj = []
            for k, sel in todo:
                g = copy.copy(sel["geom"])
                g.D = list(g.D)
                g.D[c] = g.D[c] * f
                pj.append(dict(sel, geom=g))
            rs = run(pj, dt, TN, timepoints, prec=prec, replicas=R, nthreads=nthreads)
            Fk = []
            for (k, _), r in zip(todo, rs):
                cs = None if cot_sig is None else cot_sig[k]
                tot = contract(cot_rec[k], r["rec"], cs, r["sig"])
                rep = [contract(cot_rec[k], r["rec_rep"][q], cs, None if r["sig_rep"] is None else r["sig_rep"][q])
                       for q in range(R)]
                Fk.append((tot, rep))
            F.append(Fk)
            Dv.append([p["geom"].D[c] for p in pj])
        for m, (k, _) in enumerate(todo):
            den = Dv[0][m] - Dv[1][m]
            grad[k, c] = (F[0][m][0] - F[1][m][0]) / den
            for q in range(R):
                grad_rep[k, q, c] = (F[0][m][1][q] - F[1][m][1][q]) / den
    return grad, grad_rep
