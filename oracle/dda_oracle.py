"""FP64 CPU oracle for the DDA solve path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py may import this module, and only as the checker (or the timed CPU
baseline).  The product path never calls it.

It restates the reference's algorithm in numpy/scipy, function by function
(reference file:line in each docstring):

  * geometry order, Si/Ci tables  python_bindings/dda_model/__init__.py:8-20,136-140
  * material nudge, K             src/AProductCore.cpp:25-33
  * Green's tensor lattice sum    src/AProductCore.cpp:63-244,322-458 (C, green_oracle.c)
  * FFT matvec on the 2N-1 grid   src/AProductCore.cpp:461-515
  * A*b + al*b                    src/DDAModel.cpp:1194-1203
  * BiCGSTAB                      src/DDAModel.cpp:899-1009
  * parameter mapping             src/DDAModel.cpp:37-64,163-247
  * eps -> al (LDR polarizability) src/DDAModel.cpp:317-328,1122-1151; src/Tools.cpp:353-366
  * incident field                src/DDAModel.cpp:312-316
  * local-field output            src/DDAModel.cpp:1153-1192
  * IntegratedE objective         src/ObjDDAModel.cpp:103-222
  * gradient (FD partials +
    adjoint solve)                src/DDAModel.cpp:398-578

How each part is pinned (tests/test_oracle_*.py): the golden objective of
python_tests/test_DDAModelWrapper.py:40; the geometry fixture; the textbook
free-space dyadic Green's function; the FCD -> LDR limit with exact Si/Ci;
brute-force image sums; a dense O(n^2) matvec; numpy.linalg.solve; the
Draine-Goodman LDR closed form and its Clausius-Mossotti limit; central finite
differences of the objective.  The reference's exact BiCGSTAB trajectory and
its gradient vectors are held in no reference fixture: those two outputs are
"parity unpinned" beyond the golden scalar and the FD checks.
"""
from __future__ import annotations

import ctypes
import itertools
import math
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np
import scipy.fft
from scipy.special import sici

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

METHOD_FCD = 0
METHOD_LDR = 1
# component order (00, 01, 02, 11, 12, 22), AProductCore.cpp:78-79
SYM_INDEX = ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            subprocess.run(["make", "-s", "-C", _HERE], check=True)
        lib = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        lib.oracle_green_grid.argtypes = [
            ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
            ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
            dp, dp, ctypes.c_long, ctypes.c_double, ctypes.c_int, dp]
        lib.oracle_green_grid.restype = ctypes.c_int
        lib.oracle_green_point.argtypes = [
            ctypes.c_double, ctypes.c_double, ctypes.c_int, dp, dp, ctypes.c_long,
            ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, dp]
        lib.oracle_sici_lookup.argtypes = [dp, ctypes.c_long, ctypes.c_double, ctypes.c_double]
        lib.oracle_sici_lookup.restype = ctypes.c_double
        _LIB = lib
    return _LIB


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def c_round(x):
    """C's round(): half away from zero (Python's round() is half-to-even)."""
    return math.floor(x + 0.5) if x >= 0 else -math.floor(-x + 0.5)


# ----------------------------------------------------------------------------
# inputs
# ----------------------------------------------------------------------------
def generate_geometry(nx, ny, nz):
    """x-fastest dipole order; python_bindings/dda_model/__init__.py:8-20."""
    mesh = itertools.product(range(nz), range(ny), range(nx))
    g = np.array(list(mesh))[:, ::-1]
    return g.flatten().astype(np.int64)


def sici_tables(delta=0.1, n=1_000_000):
    """Tables exactly as the wrapper builds them (__init__.py:136-140):
    samples at linspace(delta, delta*n, n-1), i.e. spacing 0.1000001, while
    the lookup assumes abscissae (p+1)*delta."""
    pos = np.linspace(delta, delta * n, n - 1)
    si, ci = sici(pos)
    return np.ascontiguousarray(si), np.ascontiguousarray(ci), delta


def sici_lookup(table, delta, x):
    """SiCi.cpp:12-44 (C restatement, called through ctypes)."""
    t = np.ascontiguousarray(table, dtype=np.float64)
    return _lib().oracle_sici_lookup(_dptr(t), t.size, delta, float(x))


def adjust_material(material, nback):
    """AProductCore.cpp:25-33: m /= nback^2, then +0.01 if |m| < 1.01."""
    out = []
    for m in material:
        m = complex(m) / (nback * nback)
        if abs(m) < 1.01:
            m = m + 0.01
        out.append(m)
    return np.array(out, dtype=np.complex128)


def wavenumber(lam, nback):
    """AProductCore.cpp:25."""
    return 2.0 * math.pi / (lam / nback)


def get_alpha(K, d, eps, n_e0, n_k):
    """LDR polarizability, Tools.cpp:353-366 (Draine & Goodman 1993)."""
    b1, b2, b3 = -1.891531, 0.1648469, -1.7700004
    eps = np.asarray(eps, dtype=np.complex128)
    a1 = (3.0 * d ** 3 / (4.0 * math.pi)) * (eps - 1.0) / (eps + 2.0)
    rad = (2.0 / 3.0) * (K * d) ** 3 * 1j
    S = (n_e0[0] * n_k[0]) ** 2 + (n_e0[1] * n_k[1]) ** 2 + (n_e0[2] * n_k[2]) ** 2
    den = 1.0 + (a1 / d ** 3) * ((b1 + eps * b2 + eps * b3 * S) * (K * d) ** 2 - rad)
    return a1 / den


def dal_deps(K, d, eps, n_e0, n_k):
    """d(1/alpha)/d eps in closed form: 1/alpha = 1/a1 + K^2 (b1+eps(b2+b3 S))/d
    - (2/3) i K^3, so d/d eps = -3*(4 pi/(3 d^3))/(eps-1)^2 + K^2 (b2+b3 S)/d."""
    b2, b3 = 0.1648469, -1.7700004
    S = (n_e0[0] * n_k[0]) ** 2 + (n_e0[1] * n_k[1]) ** 2 + (n_e0[2] * n_k[2]) ** 2
    c = 3.0 * d ** 3 / (4.0 * math.pi)
    eps = np.asarray(eps, dtype=np.complex128)
    return -3.0 / (c * (eps - 1.0) ** 2) + K * K * (b2 + b3 * S) / d


def interp_eps(x, material):
    """DDAModel.cpp:319-326 / 1122-1151: piecewise-linear material mix."""
    x = np.asarray(x, dtype=np.float64)
    lf = np.floor(x).astype(np.int64)
    ln = np.where(lf >= 1, lf, lf + 1)
    if np.any(lf < 0) or np.any(ln >= len(material)):
        raise ValueError("parameter outside the material table")
    return material[lf] + (x - lf) * (material[ln] - material[lf])


def incident_field(geometry, K, d, E0, n_e0, n_k):
    """DDAModel.cpp:312-316: E0 e exp(i K d (k . g))."""
    g = np.asarray(geometry, dtype=np.float64).reshape(-1, 3)
    ph = K * d * (n_k[0] * g[:, 0] + n_k[1] * g[:, 1] + n_k[2] * g[:, 2])
    e = np.cos(ph) + 1j * np.sin(ph)
    out = np.empty((g.shape[0], 3), dtype=np.complex128)
    for c in range(3):
        out[:, c] = E0 * n_e0[c] * e
    return out.reshape(-1)


# ----------------------------------------------------------------------------
# parameter mapping, DDAModel.cpp:37-64,163-247
# ----------------------------------------------------------------------------
def parameter_mapping(geometry, N, Nz, symmetry, symaxis):
    geometry = np.asarray(geometry, dtype=np.int64)
    if symmetry == "None":
        dividesym = 1
    elif symmetry == "4fold":
        dividesym = 4
    else:
        raise RuntimeError("Caught an unknown exception!")
    nfpara = int(geometry.size // 3 // Nz // dividesym)
    fcur = {}
    pos = 0
    gpara = np.zeros(N, dtype=np.int64)
    for i in range(N):
        xy = (int(geometry[3 * i]), int(geometry[3 * i + 1]))
        if xy not in fcur:
            if symmetry == "None":
                fcur[xy] = pos
                pos += 1
            else:
                x, y = xy
                s1 = (int(c_round(2 * symaxis[0] - x)), y)
                s2 = (int(c_round(2 * symaxis[0] - x)), int(c_round(2 * symaxis[1] - y)))
                s3 = (x, int(c_round(2 * symaxis[1] - y)))
                for s in (s1, s2, s3):
                    if s in fcur:
                        fcur[xy] = fcur[s]
                        break
                else:
                    fcur[xy] = pos
                    pos += 1
        gpara[i] = fcur[xy]
    if gpara.max(initial=-1) >= nfpara:
        raise RuntimeError("Caught an unknown exception!")
    p2g = [[] for _ in range(nfpara)]
    for i in range(N):
        p2g[gpara[i]].append(i)
    return nfpara, gpara, p2g


# ----------------------------------------------------------------------------
# Green's tensor and the FFT matvec on the reference's (2N-1)^3 grid
# ----------------------------------------------------------------------------
def green_offsets(nx, ny, nz, d, K, method, maxm, maxn, lm_nm, ln_nm,
                  si=None, ci=None, delta=0.1, octant_only=False):
    """Lattice-summed, damped interaction tensor on the centred offset grid
    [2nx-1][2ny-1][2nz-1][6] (complex128).  AProductCore.cpp:63-244."""
    if si is None:
        si = np.zeros(2)
        ci = np.zeros(2)
    si = np.ascontiguousarray(si, dtype=np.float64)
    ci = np.ascontiguousarray(ci, dtype=np.float64)
    out = np.zeros((2 * nx - 1, 2 * ny - 1, 2 * nz - 1, 6, 2), dtype=np.float64)
    _lib().oracle_green_grid(nx, ny, nz, d, K, method, maxm, maxn, lm_nm, ln_nm,
                             _dptr(si), _dptr(ci), si.size, delta, int(octant_only), _dptr(out))
    return out[..., 0] + 1j * out[..., 1]


def green_point(K, d, method, x, y, z, si=None, ci=None, delta=0.1):
    """One undamped, unsummed interaction tensor as a 3x3 complex matrix."""
    if si is None:
        si = np.zeros(2)
        ci = np.zeros(2)
    si = np.ascontiguousarray(si, dtype=np.float64)
    ci = np.ascontiguousarray(ci, dtype=np.float64)
    o = np.zeros(12)
    _lib().oracle_green_point(K, d, method, _dptr(si), _dptr(ci), si.size, delta,
                              x, y, z, _dptr(o))
    v = o[0::2] + 1j * o[1::2]
    m = np.zeros((3, 3), dtype=np.complex128)
    for c, (i, j) in enumerate(SYM_INDEX):
        m[i, j] = v[c]
        m[j, i] = v[c]
    return m


def wrap_offsets(cent, shape=None):
    """Place the centred offset grid on a wrapped FFT grid (offset o at o mod L).
    With shape=None the reference's 2N-1 grid is used (AProductCore.cpp:91-244
    stores negative offset o at o + 2N-1)."""
    gx, gy, gz = cent.shape[:3]
    nx, ny, nz = (gx + 1) // 2, (gy + 1) // 2, (gz + 1) // 2
    if shape is None:
        shape = (gx, gy, gz)
    out = np.zeros(tuple(shape) + cent.shape[3:], dtype=cent.dtype)
    ix = np.arange(-(nx - 1), nx) % shape[0]
    iy = np.arange(-(ny - 1), ny) % shape[1]
    iz = np.arange(-(nz - 1), nz) % shape[2]
    out[np.ix_(ix, iy, iz)] = cent
    return out


class FFTOperator:
    """A*b without the al*b term (AProductCore::Aproduct, AProductCore.cpp:461-515).
    The Green's spectrum is built once (AProductCore.cpp:262-279)."""

    def __init__(self, cent, geometry, nx, ny, nz, fft_shape=None, workers=1):
        self.grid = tuple(fft_shape) if fft_shape is not None else (2 * nx - 1, 2 * ny - 1, 2 * nz - 1)
        self.workers = workers
        wrapped = wrap_offsets(cent, self.grid)
        self.ahat = scipy.fft.fftn(wrapped, axes=(0, 1, 2), workers=workers)  # [gx,gy,gz,6]
        g = np.asarray(geometry, dtype=np.int64).reshape(-1, 3)
        self.ix, self.iy, self.iz = g[:, 0], g[:, 1], g[:, 2]
        self.n = g.shape[0]

    def apply(self, b):
        b = np.asarray(b, dtype=np.complex128).reshape(-1, 3)
        grid = np.zeros(self.grid + (3,), dtype=np.complex128)
        grid[self.ix, self.iy, self.iz, :] = b
        bh = scipy.fft.fftn(grid, axes=(0, 1, 2), workers=self.workers)
        a = self.ahat
        conv = np.empty_like(bh)
        conv[..., 0] = a[..., 0] * bh[..., 0] + a[..., 1] * bh[..., 1] + a[..., 2] * bh[..., 2]
        conv[..., 1] = a[..., 1] * bh[..., 0] + a[..., 3] * bh[..., 1] + a[..., 4] * bh[..., 2]
        conv[..., 2] = a[..., 2] * bh[..., 0] + a[..., 4] * bh[..., 1] + a[..., 5] * bh[..., 2]
        y = scipy.fft.ifftn(conv, axes=(0, 1, 2), workers=self.workers)  # includes the 1/NFFT
        return y[self.ix, self.iy, self.iz, :].reshape(-1)


def dense_interaction_matrix(cent, geometry):
    """O(n^2) matrix of the same operator for tiny grids: block (i,j) is
    A(g_i - g_j)."""
    g = np.asarray(geometry, dtype=np.int64).reshape(-1, 3)
    n = g.shape[0]
    gx, gy, gz = cent.shape[:3]
    nx, ny, nz = (gx + 1) // 2, (gy + 1) // 2, (gz + 1) // 2
    M = np.zeros((3 * n, 3 * n), dtype=np.complex128)
    for i in range(n):
        for j in range(n):
            o = g[i] - g[j]
            v = cent[o[0] + nx - 1, o[1] + ny - 1, o[2] + nz - 1]
            for c, (a, b) in enumerate(SYM_INDEX):
                M[3 * i + a, 3 * j + b] = v[c]
                M[3 * i + b, 3 * j + a] = v[c]
    return M


# ----------------------------------------------------------------------------
# BiCGSTAB, DDAModel.cpp:899-1009 (dot = sum conj(a) b, as Eigen's a.dot(b))
# ----------------------------------------------------------------------------
@dataclass
class SolveInfo:
    iterations: int = 0
    error: float = 0.0
    converged: bool = False
    breakdown: bool = False
    matvecs: int = 0


def bicgstab_reference(apply, E, P0, maxit, tol, history=None):
    """BiCGSTAB of DDAModel.cpp:899-1009; history (a list) receives the
    relative recursive residual after every loop step."""
    P = np.array(P0, dtype=np.complex128, copy=True)
    info = SolveInfo()
    nE = np.linalg.norm(E)
    r = E - apply(P)
    r0 = r.copy()
    p = r.copy()
    v = apply(p)
    alpha = np.vdot(r0, r) / np.vdot(r0, v)
    t = r - alpha * v
    s = apply(t)
    zeta = np.vdot(s, t) / np.vdot(s, s)
    u = zeta * v
    z = zeta * r - alpha * u
    P = P + alpha * p + z
    rl = r
    r = t - zeta * s
    info.matvecs = 3
    for it in range(maxit):
        r0rl = np.vdot(r0, rl)
        if r0rl == 0:
            info.error = np.linalg.norm(r) / nE
            info.iterations = it + 1
            info.breakdown = True
            return P, info
        beta = (alpha / zeta) * np.vdot(r0, r) / r0rl
        p = r + beta * (p - u)
        v = apply(p)
        alpha = np.vdot(r0, r) / np.vdot(r0, v)
        t = r - alpha * v
        s = apply(t)
        zeta = np.vdot(s, t) / np.vdot(s, s)
        u = zeta * v
        z = zeta * r - alpha * u
        P = P + alpha * p + z
        rl = r
        r = t - zeta * s
        info.matvecs += 2
        if history is not None:
            history.append(float(np.linalg.norm(r) / nE))
        if np.linalg.norm(r) / nE <= tol:
            info.error = np.linalg.norm(r) / nE
            info.iterations = it + 1
            info.converged = True
            return P, info
    info.iterations = maxit
    info.error = np.linalg.norm(r) / nE
    return P, info


# ----------------------------------------------------------------------------
# objective, ObjDDAModel.cpp:103-222
# ----------------------------------------------------------------------------
def objective_region(geometry, objpara):
    """Dipoles counted by ObjIntegratedEDDAModel::SingleResponse.  The
    predicate `xMin <= g <= xMax` parses as `(xMin <= g) <= xMax`
    (ObjDDAModel.cpp:135), reproduced here."""
    g = np.asarray(geometry, dtype=np.int64).reshape(-1, 3)
    lo = [int(c_round(objpara[1])), int(c_round(objpara[3])), int(c_round(objpara[5]))]
    hi = [int(c_round(objpara[2])), int(c_round(objpara[4])), int(c_round(objpara[6]))]
    ok = np.ones(g.shape[0], dtype=bool)
    for a in range(3):
        ok &= (lo[a] <= g[:, a]).astype(np.int64) <= hi[a]
    return ok


def objective_integrated(al, P, region, pownum):
    w = np.abs(np.asarray(al) * np.asarray(P)).reshape(-1, 3)
    return math.log(float(np.sum(w[region] ** pownum)))


# ----------------------------------------------------------------------------
# the model (mirrors the _dda_model.DDAModel constructor and methods)
# ----------------------------------------------------------------------------
class OracleDDAModel:
    """CPU FP64 restatement of DDAModel with the pybind11 constructor's
    argument list (python_bindings/PythonBindings.cpp:19-59)."""

    def __init__(self, betaMin, betaMax, ita, betaType, filterIterations, filterRadii,
                 useFilter, symmetryType, symmetryAxis, isPeriodic, objectiveName, objPara,
                 geometry, inputDielectrics, num_x, num_y, num_z, num_total, n_K, E0, n_E0,
                 lam, material, nback, MAXm, MAXn, Lm, Ln, matrixAMethodName, d,
                 sineIntegralValues, cosineIntegralValues, integralDelta, verbose,
                 fft_shape=None, workers=1, octant_only=True):
        geometry = np.asarray(geometry, dtype=np.int64)
        inputDielectrics = np.asarray(inputDielectrics, dtype=np.float64)
        self.N, self.Nx, self.Ny, self.Nz = int(num_total), int(num_x), int(num_y), int(num_z)
        self.geometry = geometry.copy()
        self.d = float(d)
        self.lam = float(lam)
        self.nK = np.asarray(n_K, dtype=np.float64)
        self.nE0 = np.asarray(n_E0, dtype=np.float64)
        self.E0 = float(E0)
        self.nfpara, self.gpara, self.p2g = parameter_mapping(
            geometry, self.N, self.Nz, symmetryType, list(symmetryAxis))
        if geometry.size != inputDielectrics.size:
            raise RuntimeError("Caught an unknown exception!")
        inmap = {}
        g3 = geometry.reshape(-1, 3)
        for i in range(geometry.size // 3):
            inmap.setdefault(tuple(int(v) for v in g3[i]), inputDielectrics[3 * i])
        self.parameters = np.array(
            [inmap[tuple(int(v) for v in g3[self.p2g[i][0]])] for i in range(self.nfpara)],
            dtype=np.float64)
        self.K = wavenumber(self.lam, float(nback))
        self.material = adjust_material(material, float(nback))
        if matrixAMethodName == "FCD":
            method = METHOD_FCD
        elif matrixAMethodName == "LDR":
            method = METHOD_LDR
        else:
            raise ValueError('"%s" is not a valid polarizabiliy calculation method.' % matrixAMethodName)
        self.cent = green_offsets(self.Nx, self.Ny, self.Nz, self.d, self.K, method,
                                  int(MAXm), int(MAXn), float(Lm) * self.d, float(Ln) * self.d,
                                  sineIntegralValues, cosineIntegralValues, float(integralDelta),
                                  octant_only=octant_only)
        self.op = FFTOperator(self.cent, geometry, self.Nx, self.Ny, self.Nz,
                              fft_shape=fft_shape, workers=workers)
        self.objpara = list(objPara)
        self.pownum = int(c_round(objPara[0]))
        self.region = objective_region(geometry, objPara)
        self.Einc = incident_field(geometry, self.K, self.d, self.E0, self.nE0, self.nK)
        self.E = self.Einc.copy()
        self.P = np.zeros(3 * self.N, dtype=np.complex128)
        self.EResult = np.zeros(3 * self.N, dtype=np.complex128)
        self._update_alpha()
        self.info = SolveInfo()

    # DDAModel.cpp:747-772,1122-1151
    def _update_alpha(self):
        self.diel_old = np.repeat(self.parameters[self.gpara], 3)
        self.diel = interp_eps(self.diel_old, self.material)
        self.al = 1.0 / get_alpha(self.K, self.d, self.diel, self.nE0, self.nK)

    def apply(self, b):
        """Aproductwithalb, DDAModel.cpp:1194-1203."""
        return self.op.apply(b) + self.al * b

    def setParameters(self, values):
        values = np.asarray(values, dtype=np.float64)
        if values.size != self.nfpara:
            raise RuntimeError("Caught an unknown exception!")
        self.parameters = values.copy()
        self._update_alpha()

    def getParameters(self):
        return self.parameters.copy()

    def getDielectrics(self):
        return self.diel_old.copy()

    def getElectricField(self):
        return self.EResult.copy()

    def solveElectricField(self, originPolarization, maxit, tol):
        """DDAModel.cpp:874-885, with update_E_in_structure (:1185-1192)."""
        self.P, self.info = bicgstab_reference(self.apply, self.E,
                                               np.asarray(originPolarization, dtype=np.complex128),
                                               int(maxit), float(tol))
        self.EResult = self.al * self.P / ((2.0 + self.diel) / 3.0)

    def calculateObjective(self):
        return objective_integrated(self.al, self.P, self.region, self.pownum)

    # --- gradients -------------------------------------------------------
    def _al_of(self, x):
        eps = interp_eps(np.asarray(x, dtype=np.float64), self.material)
        return 1.0 / get_alpha(self.K, self.d, eps, self.nE0, self.nK)

    def fd_partials(self, eps, origin):
        """Vectorised restatement of devx_and_Adevxp and devp
        (DDAModel.cpp:471-578): every FD quotient uses one perturbed sum with
        the other terms held fixed, which is what the reference's O(1)
        incremental SingleResponse updates compute, minus their rounding drift."""
        p = self.pownum
        w = np.abs(self.al * self.P) ** p
        reg3 = np.repeat(self.region, 3)
        Eint = float(np.sum(w[reg3]))
        sign = np.where(self.parameters >= eps, -1.0, 1.0)
        al_pert = self._al_of(self.parameters + sign * eps)  # per class
        devx = np.zeros(self.nfpara)
        adevxp = np.zeros(3 * self.N, dtype=np.complex128)
        for i in range(self.nfpara):
            dips = np.asarray(self.p2g[i], dtype=np.int64)
            idx = (3 * dips[:, None] + np.arange(3)[None, :]).reshape(-1)
            alp = al_pert[i]
            delta = float(np.sum((np.abs(alp * self.P[idx]) ** p - w[idx])[reg3[idx]]))
            devx[i] = (math.log(Eint + delta) - origin) / (sign[i] * eps)
            adevxp[idx] = (alp - self.al[idx]) / (sign[i] * eps)
        adevxp = adevxp * self.P
        d_re = np.abs(self.al * (self.P + eps)) ** p - w
        d_im = np.abs(self.al * (self.P + 1j * eps)) ** p - w
        d_re = np.where(reg3, d_re, 0.0)
        d_im = np.where(reg3, d_im, 0.0)
        devp = (np.log(Eint + d_re) - origin) / eps + (np.log(Eint + d_im) - origin) / (1j * eps)
        return devx, adevxp, devp

    def calculateGradients(self, epsilon_partial, originalObjValue, maxit, tol):
        """DDAModel.cpp:398-469.  Side effects kept: P <- lambda, E <- E_inc."""
        devx, adevxp, devp = self.fd_partials(float(epsilon_partial), float(originalObjValue))
        lam, info = bicgstab_reference(self.apply, devp, np.zeros_like(devp), int(maxit), float(tol))
        self.adjoint_info = info
        self.P = lam
        self.E = self.Einc.copy()
        mult = np.zeros(self.nfpara, dtype=np.complex128)
        prod = lam * adevxp
        for i in range(self.nfpara):
            dips = np.asarray(self.p2g[i], dtype=np.int64)
            mult[i] = np.sum(prod[(3 * dips[:, None] + np.arange(3)[None, :]).reshape(-1)])
        return devx - mult.real

    def exact_gradient(self, tol=1e-12, maxit=100000):
        """Exact derivative of the objective w.r.t. the free parameters by the
        adjoint method (closed-form partials; same linear algebra as the
        reference gradient but without finite differences).  Requires a
        converged forward P."""
        p = self.pownum
        w_c = self.al * self.P
        aw = np.abs(w_c)
        reg3 = np.repeat(self.region, 3)
        Eint = float(np.sum((aw ** p)[reg3]))
        # d|w|^p / dP (Wirtinger, times 2) = p |w|^(p-2) conj(w) al
        with np.errstate(divide="ignore", invalid="ignore"):
            g = np.where(aw > 0, p * aw ** (p - 2) * np.conj(w_c) * self.al, 0.0)
        devp = np.where(reg3, g, 0.0) / Eint
        dal = dal_deps(self.K, self.d, self.diel, self.nE0, self.nK)
        deps = self.material[1] - self.material[0]
        dal_drho = dal * np.where(self.diel_old <= 1.0, deps, 0.0)
        # explicit part: d|al P|^p / d rho = p |w|^(p-2) Re(conj(w) P dal)
        with np.errstate(divide="ignore", invalid="ignore"):
            dexp = np.where(aw > 0, p * aw ** (p - 2) * np.real(np.conj(w_c) * self.P * dal_drho), 0.0)
        dexp = np.where(reg3, dexp, 0.0) / Eint
        lam, info = bicgstab_reference(self.apply, devp, np.zeros_like(devp), maxit, tol)
        contrib = dexp - np.real(lam * dal_drho * self.P)
        grad = np.zeros(self.nfpara)
        for i in range(self.nfpara):
            dips = np.asarray(self.p2g[i], dtype=np.int64)
            grad[i] = np.sum(contrib[(3 * dips[:, None] + np.arange(3)[None, :]).reshape(-1)])
        return grad


class OracleFullBox:
    """A full Nx*Ny*Nz box (x-fastest dipoles) with one density per voxel: the
    same operator, polarizability, incident field and BiCGSTAB as the model
    above, without the parameter classes (eps from rho per voxel,
    DDAModel.cpp:319-328).  Densities are clipped to [0, 1].  Objectives:
    IntegratedE over all dipoles (ObjDDAModel.cpp:130-212) and the mean |E|^2
    over external points, E(r) = E_inc(r) - sum_j A(r - r_j) P_j with the
    undamped single-image interaction (DDAModel::solve_E, DDAModel.cpp:1153-1183,
    incident field E0 e exp(i K k.r) in every component)."""

    def __init__(self, shape, d, lam, rho, n_k=(0, 0, 1), n_e0=(1, 0, 0), e0=1.0,
                 material=(1.0 + 0j, 5.96 + 3.89j), nback=1.0, method="LDR", images=0, period=(0, 0),
                 tables=None, workers=1):
        self.shape = tuple(shape)
        nx, ny, nz = shape
        self.N = nx * ny * nz
        self.d, self.lam = float(d), float(lam)
        self.nK, self.nE0, self.E0 = np.asarray(n_k, float), np.asarray(n_e0, float), float(e0)
        self.K = wavenumber(self.lam, nback)
        self.material = adjust_material(material, nback)
        self.method = METHOD_FCD if method == "FCD" else METHOD_LDR
        si, ci, delta = tables if tables is not None else (np.zeros(2), np.zeros(2), 0.1)
        self.tables = (si, ci, delta)
        self.geometry = generate_geometry(nx, ny, nz)
        self.cent = green_offsets(nx, ny, nz, self.d, self.K, self.method, images, images,
                                  period[0] * self.d, period[1] * self.d, si, ci, delta, octant_only=True)
        self.op = FFTOperator(self.cent, self.geometry, nx, ny, nz, workers=workers)
        self.Einc = incident_field(self.geometry, self.K, self.d, self.E0, self.nE0, self.nK)
        self.set_density(rho)

    def set_density(self, rho):
        self.rho = np.clip(np.asarray(rho, float).reshape(-1), 0.0, 1.0)
        self.eps = interp_eps(self.rho, self.material)
        self.al = np.repeat(1.0 / get_alpha(self.K, self.d, self.eps, self.nE0, self.nK), 3)

    def apply(self, b):
        return self.op.apply(b) + self.al * b

    def solve(self, tol=1e-12, maxit=100000, rhs=None):
        rhs = self.Einc if rhs is None else rhs
        P, info = bicgstab_reference(self.apply, rhs, np.zeros(3 * self.N, complex), maxit, tol)
        if rhs is self.Einc:
            self.P, self.info = P, info
        return P, info

    def field(self):
        return self.al * self.P / ((2.0 + np.repeat(self.eps, 3)) / 3.0)

    def objective_integrated(self, power=2):
        return math.log(float(np.sum(np.abs(self.al * self.P) ** power)))

    def _pair(self, r):
        si, ci, delta = self.tables
        return green_point(self.K, self.d, self.method, *r, si=si, ci=ci, delta=delta)

    def target_fields(self, targets):
        g = self.geometry.reshape(-1, 3) * self.d
        out = []
        for t in np.asarray(targets, float).reshape(-1, 3):
            ph = self.K * float(self.nK @ t)
            e = self.E0 * self.nE0 * complex(math.cos(ph), math.sin(ph))
            s = np.zeros(3, complex)
            for j in range(self.N):
                s += self._pair(t - g[j]) @ self.P[3 * j:3 * j + 3]
            out.append(e - s)
        return np.array(out)

    def target_fields_ldr(self, targets):
        """target_fields for the LDR interaction, vectorised over the dipoles
        (numpy): the same undamped single-image tensor as green_point
        (AProductCore.cpp:351-386), E(t) = E_inc(t) - sum_j A(t - r_j) P_j.
        tests/ check it against the loop version."""
        if self.method != METHOD_LDR:
            raise ValueError("target_fields_ldr evaluates the LDR interaction only")
        g = self.geometry.reshape(-1, 3).astype(np.float64) * self.d
        P = self.P.reshape(-1, 3)
        K = self.K
        out = []
        for t in np.asarray(targets, float).reshape(-1, 3):
            R = t[None, :] - g
            xx, yy, zz = R[:, 0] ** 2, R[:, 1] ** 2, R[:, 2] ** 2
            r2 = xx + yy + zz
            r = np.sqrt(r2)
            with np.errstate(divide="ignore", invalid="ignore"):
                pre = np.where(r > 0, np.exp(1j * K * r) / (r2 * r), 0.0)
                c2 = np.where(r > 0, (1.0 - 1j * K * r) / r2, 0.0)
            c3 = K * K - 3.0 * c2
            a00 = pre * (c2 * (r2 - 3 * xx) - K * K * (yy + zz))
            a11 = pre * (c2 * (r2 - 3 * yy) - K * K * (zz + xx))
            a22 = pre * (c2 * (r2 - 3 * zz) - K * K * (xx + yy))
            a01 = pre * c3 * R[:, 0] * R[:, 1]
            a02 = pre * c3 * R[:, 0] * R[:, 2]
            a12 = pre * c3 * R[:, 1] * R[:, 2]
            s = np.array([np.sum(a00 * P[:, 0] + a01 * P[:, 1] + a02 * P[:, 2]),
                          np.sum(a01 * P[:, 0] + a11 * P[:, 1] + a12 * P[:, 2]),
                          np.sum(a02 * P[:, 0] + a12 * P[:, 1] + a22 * P[:, 2])])
            ph = K * float(self.nK @ t)
            out.append(self.E0 * self.nE0 * complex(math.cos(ph), math.sin(ph)) - s)
        return np.array(out)

    def objective_targets(self, targets):
        E = self.target_fields(targets)
        return float(np.sum(np.abs(E) ** 2)) / E.shape[0]

    def gradient(self, kind="integrated", power=2, targets=None, tol=1e-12):
        """d objective / d rho per voxel: closed-form partials, one adjoint solve."""
        w = self.al * self.P
        dal = np.repeat(dal_deps(self.K, self.d, self.eps, self.nE0, self.nK) * (self.material[1] - self.material[0]), 3)
        explicit = np.zeros(3 * self.N)
        if kind == "integrated":
            eint = float(np.sum(np.abs(w) ** power))
            aw = np.abs(w)
            with np.errstate(divide="ignore", invalid="ignore"):
                s = np.where(aw > 0, power * aw ** (power - 2), 0.0) if power != 2 else np.full_like(aw, 2.0)
            devp = s * np.conj(w) * self.al / eint
            explicit = s * np.real(np.conj(w) * dal * self.P) / eint
        else:
            E = self.target_fields(targets)
            T = E.shape[0]
            g = self.geometry.reshape(-1, 3) * self.d
            devp = np.zeros(3 * self.N, complex)
            for j in range(self.N):
                acc = np.zeros(3, complex)
                for t, tt in enumerate(np.asarray(targets, float).reshape(-1, 3)):
                    acc += self._pair(tt - g[j]) @ np.conj(E[t])
                devp[3 * j:3 * j + 3] = -2.0 / T * acc
        lam, _ = self.solve(tol=tol, rhs=devp)
        contrib = explicit - np.real(lam * dal * self.P)
        return contrib.reshape(-1, 3).sum(axis=1)


def reference_gradient_sequential(model, eps, origin):
    """Literal restatement of devx_and_Adevxp and devp with the reference's
    incremental E_int bookkeeping (DDAModel.cpp:471-578, ObjDDAModel.cpp:130-157),
    in pure Python loops: for small cases only (tests check the vectorised
    fd_partials against it)."""
    p = model.pownum
    al = model.al.copy()
    P = model.P.copy()
    region = model.region
    state = {"E": 0.0}

    def single(idx, deduction):
        if not region[idx]:
            return
        s = 0.0
        for c in range(3):
            s_c = abs(al[3 * idx + c] * P[3 * idx + c]) ** p
            state["E"] = state["E"] - s_c if deduction else state["E"] + s_c
        return s

    for idx in range(model.N):
        single(idx, False)
    devx = np.zeros(model.nfpara)
    adevxp = np.zeros(3 * model.N, dtype=np.complex128)
    diel_old = model.diel_old.copy()
    for i in range(model.nfpara):
        orig = model.parameters[i]
        sign = -1 if orig >= eps else 1
        tmp = orig + sign * eps
        for pos in model.p2g[i]:
            a0 = al[3 * pos]
            single(pos, True)
            diel_old[3 * pos:3 * pos + 3] = tmp
            new = model._al_of(np.array([tmp]))[0]
            al[3 * pos:3 * pos + 3] = new
            single(pos, False)
            adevxp[3 * pos:3 * pos + 3] = (al[3 * pos] - a0) / (sign * eps)
        devx[i] = (math.log(state["E"]) - origin) / (sign * eps)
        for pos in model.p2g[i]:
            single(pos, True)
            diel_old[3 * pos:3 * pos + 3] = orig
            al[3 * pos:3 * pos + 3] = model._al_of(np.array([orig]))[0]
            single(pos, False)
    adevxp = adevxp * P
    devp = np.zeros(3 * model.N, dtype=np.complex128)
    for i in range(3 * model.N):
        pos = i // 3
        single(pos, True)
        P[i] = P[i] + eps
        single(pos, False)
        devp[i] += (math.log(state["E"]) - origin) / eps
        single(pos, True)
        P[i] = P[i] - eps
        P[i] = P[i] + 1j * eps
        single(pos, False)
        devp[i] += (math.log(state["E"]) - origin) / (1j * eps)
        single(pos, True)
        P[i] = P[i] - 1j * eps
        single(pos, False)
    return devx, adevxp, devp


def wrapper_args(num_pixels_xyz, pixel_size_nm, initial_parameter_values, dielectric_constants,
                 light_direction, light_polarization, light_wavelength_nm, light_amplitude=1.0,
                 symmetry_axes=(), symmetry_method="4fold", symmetry_is_periodic=True,
                 integral_power=2, integral_xbounds=None, integral_ybounds=None,
                 integral_zbounds=None, background_refractive_index=None,
                 objective_name="IntegratedE", apc_max_m=50, apc_max_n=50, apc_l_m=None,
                 apc_l_n=None, apc_method_name="FCD", verbose=False, tables=None):
    """The positional DDAModel argument list DDAModelWrapper builds
    (python_bindings/dda_model/__init__.py:99-160)."""
    nx, ny, nz = num_pixels_xyz
    geometry = generate_geometry(nx, ny, nz)
    integral_xbounds = integral_xbounds or [0.0, float(nx) - 1]
    integral_ybounds = integral_ybounds or [0.0, float(ny) - 1]
    integral_zbounds = integral_zbounds or [0.0, float(nz) - 1]
    cfg = [integral_power] + list(integral_xbounds) + list(integral_ybounds) + list(integral_zbounds) + [0.95, 50.0]
    apc_l_n = apc_l_n or nx
    apc_l_m = apc_l_m or ny
    si, ci, delta = tables if tables is not None else sici_tables()
    if not background_refractive_index:
        background_refractive_index = float(np.sqrt(np.real(dielectric_constants[0])))
    return [0.0, 50.0, 0.5, "piecewise", [100], [2.0], False,
            symmetry_method, list(symmetry_axes), symmetry_is_periodic,
            objective_name, cfg, geometry, initial_parameter_values,
            nx, ny, nz, int(np.prod(num_pixels_xyz)),
            list(light_direction), light_amplitude, list(light_polarization),
            light_wavelength_nm, list(dielectric_constants), background_refractive_index,
            apc_max_m, apc_max_n, apc_l_m, apc_l_n, apc_method_name, pixel_size_nm,
            si, ci, delta, verbose]
