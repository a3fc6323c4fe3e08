"""ctypes wrapper of the CPU oracle (oracle/libce_oracle.so) -- test infrastructure only.

Imported by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg; the product never
imports it.  The oracle restates the reference chain in FP64 (see ce_oracle.cpp for the file:line
map of the reference).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# CE_ORACLE_LIB names another build of the oracle (tests/test_oracle_mutants.py loads deliberately
# broken copies through it to show that the pinning tests catch each mistake)
LIB_PATH = os.environ.get("CE_ORACLE_LIB") or os.path.join(HERE, "libce_oracle.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/libce_oracle.so is not built: run python oracle/build.py")
        _lib = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        _lib.ora_evaluate.argtypes = [vp, vp, vp, vp, vp, vp, ctypes.c_int, ctypes.c_int, vp, vp, vp, vp, vp,
                                      ctypes.c_char_p, ctypes.c_int]
        _lib.ora_evaluate.restype = ctypes.c_int
        _lib.ora_derive.argtypes = [vp, vp, ctypes.c_char_p, ctypes.c_int]
        _lib.ora_decode.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, ctypes.c_int, vp, vp, ctypes.c_char_p,
                                    ctypes.c_int]
        _lib.ora_interp.argtypes = [vp, vp, ctypes.c_int, vp, ctypes.c_int, ctypes.c_double]
        _lib.ora_interp.restype = ctypes.c_double
        _lib.ora_eigvals.argtypes = [vp, ctypes.c_int, vp, vp]
        _lib.ora_eigvals_events.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, vp, vp]
        _lib.ora_ldlt_solve.argtypes = [vp, ctypes.c_int, vp, vp]
        _lib.ora_ldlt_solve.restype = None
        _lib.ora_shifted_solve.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, vp, vp, vp]
        _lib.ora_detmath.argtypes = [ctypes.c_int, vp, vp, vp, ctypes.c_int]
        _lib.ora_grid.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp, ctypes.c_char_p, ctypes.c_int]
        _lib.ora_max_threads.restype = ctypes.c_int
        _lib.ora_simulate_kalman.argtypes = [vp, vp, vp, ctypes.c_int, vp, ctypes.c_char_p, ctypes.c_int]
        _lib.ora_simulate_kalman.restype = ctypes.c_int
        _lib.ora_evaluate_ref32.argtypes = [vp, vp, vp, vp, vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_char_p,
                                            ctypes.c_int]
        _lib.ora_evaluate_ref32.restype = ctypes.c_int
        _lib.ora_ldlt_solve32.argtypes = [vp, ctypes.c_int, vp, vp]
        _lib.ora_ldlt_solve32.restype = None
        _lib.ora_shifted_solve32.argtypes = [vp, ctypes.c_int, ctypes.c_float, ctypes.c_float, vp, vp, vp]
        _lib.ora_shifted_solve32.restype = ctypes.c_int
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def evaluate(problem, rnd=0, gains=None, info=None, threads=0, diag=False):
    """FP64 chain over the candidate rows of one round.  Returns (info32, info64, diag dict)."""
    cfg = problem.cfg_struct()
    plant = problem.plant_struct(rnd)
    g = np.ascontiguousarray(problem.gains if gains is None else gains, dtype=np.float32)
    inf = np.ascontiguousarray(problem.info if info is None else info, dtype=np.float32)
    n = g.shape[0]
    out32 = np.zeros_like(inf)
    out64 = np.zeros(inf.shape, dtype=np.float64)
    dg = {}
    if diag:
        d = problem.derived()
        nt = d.total_dim
        dg = dict(y=np.zeros((n, cfg.time)), lina=np.zeros((n, nt * nt)), eig=np.zeros((n, 2 * nt)),
                  ata=np.zeros((n, 4 * nt * nt)), atb=np.zeros((n, 2 * nt)))
    err = ctypes.create_string_buffer(1024)
    rc = lib().ora_evaluate(ctypes.byref(cfg), ctypes.byref(plant), _p(g), _p(inf), _p(out32), _p(out64), n,
                            int(threads), _p(dg.get("y")), _p(dg.get("lina")), _p(dg.get("eig")),
                            _p(dg.get("ata")), _p(dg.get("atb")), err, len(err))
    if rc != 0:
        raise ValueError(err.value.decode())
    return out32, out64, dg


def evaluate_ref32(problem, rnd=0, gains=None, info=None, threads=0, want_y=False):
    """The reference's own precision map (float simulation with hi/lo state, float eigensetup, double
    eigensolve stored as float, float modal fit and LDLT, complex<float> margin sweep) over the
    candidate rows of one round.  Returns (info32, y32 or None)."""
    cfg = problem.cfg_struct()
    plant = problem.plant_struct(rnd)
    g = np.ascontiguousarray(problem.gains if gains is None else gains, dtype=np.float32)
    inf = np.ascontiguousarray(problem.info if info is None else info, dtype=np.float32)
    n = g.shape[0]
    out = np.zeros_like(inf)
    y = np.zeros((n, cfg.time), dtype=np.float32) if want_y else None
    err = ctypes.create_string_buffer(512)
    rc = lib().ora_evaluate_ref32(ctypes.byref(cfg), ctypes.byref(plant), _p(g), _p(inf), _p(out), _p(y), n,
                                  int(threads), err, len(err))
    if rc != 0:
        raise ValueError(err.value.decode())
    return out, y


def simulate_kalman(problem, gains, rnd=0):
    """Trajectories (n x time, double) of the Kalman simulator for gains rows [K, N, R, Q] (2 ns + 2)."""
    cfg = problem.cfg_struct()
    plant = problem.plant_struct(rnd)
    g = np.ascontiguousarray(gains, dtype=np.float32)
    n = g.shape[0]
    y = np.zeros((n, cfg.time))
    err = ctypes.create_string_buffer(512)
    rc = lib().ora_simulate_kalman(ctypes.byref(cfg), ctypes.byref(plant), _p(g), n, _p(y), err, len(err))
    if rc != 0:
        raise ValueError(err.value.decode())
    return y


def eigvals(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    n = a.shape[0]
    wr = np.zeros(n)
    wi = np.zeros(n)
    st = lib().ora_eigvals(_p(a), n, _p(wr), _p(wi))
    return st, wr + 1j * wi


def eigvals_events(a, sign_block):
    """eigvals plus the flag of the solve's sign-flip events for D = diag(I_sign_block, -I)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    n = a.shape[0]
    wr = np.zeros(n)
    wi = np.zeros(n)
    ev = ctypes.c_int(0)
    st = lib().ora_eigvals_events(_p(a), n, int(sign_block), _p(wr), _p(wi), ctypes.byref(ev))
    return st, wr + 1j * wi, ev.value


def ldlt_solve(m, b):
    m = np.ascontiguousarray(m, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros_like(b)
    lib().ora_ldlt_solve(_p(m), m.shape[0], _p(b), _p(x))
    return x


def shifted_solve(a, z, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    n = a.shape[0]
    xr = np.zeros(n)
    xi = np.zeros(n)
    ok = lib().ora_shifted_solve(_p(a), n, float(np.real(z)), float(np.imag(z)), _p(b), _p(xr), _p(xi))
    return bool(ok), xr + 1j * xi


def ldlt_solve32(m, b):
    m = np.ascontiguousarray(m, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    x = np.zeros_like(b)
    lib().ora_ldlt_solve32(_p(m), m.shape[0], _p(b), _p(x))
    return x


def shifted_solve32(a, z, b):
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    n = a.shape[0]
    xr = np.zeros(n, dtype=np.float32)
    xi = np.zeros(n, dtype=np.float32)
    ok = lib().ora_shifted_solve32(_p(a), n, float(np.real(z)), float(np.imag(z)), _p(b), _p(xr), _p(xi))
    return bool(ok), xr.astype(np.float64) + 1j * xi.astype(np.float64)


def detmath(fn, x, y=None):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(x if y is None else y, dtype=np.float64)
    out = np.zeros_like(x)
    lib().ora_detmath(int(fn), _p(x), _p(y), _p(out), x.size)
    return out


def decode(info, dim, count, base, n_params):
    info = np.ascontiguousarray(info, dtype=np.int32)
    base = np.ascontiguousarray(base, dtype=np.float32)
    iout = np.zeros((max(count, 1), 4), dtype=np.int32)
    vout = np.zeros((max(count, 1), 2), dtype=np.float64)
    err = ctypes.create_string_buffer(512)
    rc = lib().ora_decode(_p(info), dim, count, _p(base), n_params, _p(iout), _p(vout), err, len(err))
    if rc != 0:
        raise ValueError(err.value.decode())
    return iout[:count], vout[:count]


def interp(iops, vops, src, seed=None):
    iops = np.ascontiguousarray(iops, dtype=np.int32)
    vops = np.ascontiguousarray(vops, dtype=np.float64)
    src = np.ascontiguousarray(src, dtype=np.float64)
    return lib().ora_interp(_p(iops), _p(vops), iops.shape[0], _p(src), 0 if seed is None else 1,
                            0.0 if seed is None else float(seed))


def grid(problem, rnd=0):
    cfg = problem.cfg_struct()
    plant = problem.plant_struct(rnd)
    F = cfg.freq
    arrs = [np.zeros(F) for _ in range(5)]
    pok = np.zeros(F, dtype=np.int32)
    alpha = ctypes.c_double()
    err = ctypes.create_string_buffer(512)
    rc = lib().ora_grid(ctypes.byref(cfg), ctypes.byref(plant), *[_p(a) for a in arrs], _p(pok), ctypes.byref(alpha),
                        err, len(err))
    if rc != 0:
        raise ValueError(err.value.decode())
    zr, zi, th, pr, pi = arrs
    return dict(z=zr + 1j * zi, theta=th, presp=pr + 1j * pi, pok=pok, alpha=alpha.value)
