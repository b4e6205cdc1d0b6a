"""CPU oracle: restatement of Helmholtz.jl's discrete Helmholtz operator.

TEST INFRASTRUCTURE ONLY.  This module is the checker for the HIP path; it is
imported by ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` and by nothing else.  The product path never calls it.

Every function restates one reference function (``src/...`` paths are relative
to the Helmholtz.jl repository) with SciPy sparse matrices built by the same
Kronecker recipe the reference uses.  Vectors are 1-D, column-major over the
nodal grid (x fastest), i.e. ``p = i + j*n1 (+ k*n1*n2)`` 0-based, the
reference's ``loc2cs`` minus one.

Pins (tests/test_oracle.py): explicit 4x4 KATs of ``dxx_mat``; ``L @ 1 == 0``;
BC=1 Laplacian == the reference's own matrix-free ``multOpNeumann!``/
``Lap2DStencil`` restated; ABL closed-form values on tiny grids; Sommerfeld
closed form; shifted-param == shift-op identity; manufactured-solution second
order convergence (the flow of ``test/testFictitiousSource2D.jl``); attenuation
equivalence (``test/AttenuationTest.jl`` flow); committed golden vectors.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp


# --------------------------------------------------------------------------
# mesh (jInv.Mesh.RegularMesh semantics: n = cells, nodes = n + 1)
# --------------------------------------------------------------------------
@dataclass
class RegularMesh:
    domain: tuple
    n: np.ndarray  # cells per axis
    h: np.ndarray = field(init=False)
    dim: int = field(init=False)

    def __post_init__(self):
        self.n = np.asarray(self.n, dtype=np.int64).ravel()
        dom = np.asarray(self.domain, dtype=np.float64).ravel()
        self.dim = len(self.n)
        self.h = np.array([(dom[2 * d + 1] - dom[2 * d]) / self.n[d] for d in range(self.dim)])

    @property
    def nodes(self):
        return self.n + 1


def get_regular_mesh(domain, ncells) -> RegularMesh:
    return RegularMesh(tuple(np.asarray(domain, dtype=float).ravel()), np.asarray(ncells))


# --------------------------------------------------------------------------
# src/PlainNodalLaplacian.jl:4-46
# --------------------------------------------------------------------------
def get_bc(order_neumann_bc: int) -> float:
    """``getBC`` (src/PlainNodalLaplacian.jl:4-15): 2 -> 2.0, 1 -> 1.0."""
    return 1.0 if order_neumann_bc == 1 else 2.0


def dxx_mat(n: int, h: float, order_neumann_bc: int = 2) -> sp.csr_matrix:
    """``dxxMat`` (src/PlainNodalLaplacian.jl:18-30): tridiagonal 1-D second
    difference with first/last rows [BC, -BC]/h^2."""
    bc = get_bc(order_neumann_bc)
    lower = -np.ones(n - 1)
    lower[n - 2] = -bc
    main = 2.0 * np.ones(n)
    main[0] = bc
    main[n - 1] = bc
    upper = -np.ones(n - 1)
    upper[0] = -bc
    return sp.diags([lower / h**2, main / h**2, upper / h**2], [-1, 0, 1], format="csr")


def get_nodal_laplacian_matrix(mesh: RegularMesh, order_neumann_bc: int = 2) -> sp.csr_matrix:
    """``getNodalLaplacianMatrix`` (src/PlainNodalLaplacian.jl:32-46):
    2-D  L = I2 (x) D1 + D2 (x) I1;  3-D  L = I3 (x) (I2 (x) D1 + D2 (x) I1) + D3 (x) I2 (x) I1."""
    nodes = mesh.nodes
    I1 = sp.identity(nodes[0], format="csr")
    I2 = sp.identity(nodes[1], format="csr")
    D1 = dxx_mat(nodes[0], mesh.h[0], order_neumann_bc)
    D2 = dxx_mat(nodes[1], mesh.h[1], order_neumann_bc)
    if mesh.dim == 2:
        return (sp.kron(I2, D1) + sp.kron(D2, I1)).tocsr()
    I3 = sp.identity(nodes[2], format="csr")
    D3 = dxx_mat(nodes[2], mesh.h[2], order_neumann_bc)
    return (sp.kron(I3, sp.kron(I2, D1) + sp.kron(D2, I1)) + sp.kron(D3, sp.kron(I2, I1))).tocsr()


# --------------------------------------------------------------------------
# spread ("HO") Laplacian and mass: src/PlainNodalLaplacian.jl:54-141,
# src/GetHelmholtz.jl:18-20,54-72
# --------------------------------------------------------------------------
def _speye(n: int) -> sp.csr_matrix:
    return sp.identity(int(n), format="csr")


def ddx_cn(n: int, h: float) -> sp.csr_matrix:
    """``ddxCN`` (src/PlainNodalLaplacian.jl:54-58): n x (n+1) cell-to-node
    difference, row c = (-1/h at node c, +1/h at node c+1)."""
    rows = np.repeat(np.arange(n), 2)
    cols = np.stack([np.arange(n), np.arange(n) + 1], axis=1).ravel()
    vals = np.tile([-1.0 / h, 1.0 / h], n)
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n + 1))


def av3term(n: int, alpha: float = 5.0 / 6.0) -> sp.csr_matrix:
    """``av3term`` (src/PlainNodalLaplacian.jl:60-66): tridiagonal average,
    diagonal alpha, off-diagonals (1-alpha)/2, end entries 1/2 + alpha/2."""
    t = (1.0 - alpha) / 2.0
    T = sp.diags([np.full(n - 1, t), np.full(n, alpha), np.full(n - 1, t)], [-1, 0, 1], format="lil")
    T[0, 0] = 0.5 + alpha / 2.0
    T[n - 1, n - 1] = 0.5 + alpha / 2.0
    return T.tocsr()


def get_nodal_spread_gradients(mesh: RegularMesh, av_func):
    """``getNodalSpreadGradients`` (src/PlainNodalLaplacian.jl:69-100): the
    nodal gradient G and its transversally averaged version Gs."""
    n = [int(v) for v in mesh.n]
    h = mesh.h
    if mesh.dim == 2:
        t = ddx_cn(n[0], h[0])
        D1, D1s = sp.kron(_speye(n[1] + 1), t), sp.kron(av_func(n[1] + 1), t)
        t = ddx_cn(n[1], h[1])
        D2, D2s = sp.kron(t, _speye(n[0] + 1)), sp.kron(t, av_func(n[0] + 1))
        return sp.vstack([D1, D2]).tocsr(), sp.vstack([D1s, D2s]).tocsr()
    I1, I2, I3 = _speye(n[0] + 1), _speye(n[1] + 1), _speye(n[2] + 1)
    t = ddx_cn(n[0], h[0])
    D1 = sp.kron(I3, sp.kron(I2, t))
    D1s = 0.5 * (sp.kron(I3, sp.kron(av_func(n[1] + 1), t)) + sp.kron(av_func(n[2] + 1), sp.kron(I2, t)))
    t = ddx_cn(n[1], h[1])
    D2 = sp.kron(I3, sp.kron(t, I1))
    D2s = 0.5 * (sp.kron(I3, sp.kron(t, av_func(n[0] + 1))) + sp.kron(av_func(n[2] + 1), sp.kron(t, I1)))
    t = ddx_cn(n[2], h[2])
    D3 = sp.kron(t, sp.kron(I2, I1))
    D3s = 0.5 * (sp.kron(t, sp.kron(I2, av_func(n[0] + 1))) + sp.kron(t, sp.kron(av_func(n[1] + 1), I1)))
    return sp.vstack([D1, D2, D3]).tocsr(), sp.vstack([D1s, D2s, D3s]).tocsr()


def get_spread_nodal_laplacian_and_mass(mesh: RegularMesh, beta):
    """``getSpreadNodalLaplacianAndMass`` (src/PlainNodalLaplacian.jl:106-141):
    Lap = G' ((1-b) Gs + b G) and the averaging mass matrix M.  2-D: scalar
    beta for both; 3-D: beta = (b_lap, b_mass) (a scalar 1 means (1, 1))."""
    n = [int(v) for v in mesh.n]
    av = lambda k: av3term(k, 0.5)
    G, Gs = get_nodal_spread_gradients(mesh, av)
    if mesh.dim == 2:
        b = float(beta)
        Lap = (G.T @ ((1 - b) * Gs + b * G)).tocsr()
        M = 0.5 * sp.kron(av3term(n[1] + 1, b), _speye(n[0] + 1)) + 0.5 * sp.kron(_speye(n[1] + 1), av3term(n[0] + 1, b))
        return Lap, M.tocsr()
    bb = np.broadcast_to(np.asarray(beta, dtype=float).ravel(), (2,))
    Lap = (G.T @ ((1 - bb[0]) * Gs + bb[0] * G)).tocsr()
    I1, I2, I3 = _speye(n[0] + 1), _speye(n[1] + 1), _speye(n[2] + 1)
    M = (1.0 / 3.0) * (sp.kron(I3, sp.kron(av3term(n[1] + 1, bb[1]), I1))
                       + sp.kron(I3, sp.kron(I2, av3term(n[0] + 1, bb[1])))
                       + sp.kron(av3term(n[2] + 1, bb[1]), sp.kron(I2, I1)))
    return Lap, M.tocsr()


def get_helmholtz_operator_ho(mesh: RegularMesh, m_nodal, omega, gamma, neumann_at_first_dim: bool,
                              sommerfeld: bool, beta=1.0) -> sp.csr_matrix:
    """``GetHelmholtzOperatorHO`` (src/GetHelmholtz.jl:54-72): H = Lap + M diag(mass)."""
    Lap, M = get_spread_nodal_laplacian_and_mass(mesh, beta)
    mass = helmholtz_mass(mesh, m_nodal, omega, gamma, neumann_at_first_dim, sommerfeld)
    return (Lap.astype(complex) + M @ sp.diags(mass)).tocsr()


def lap2d_stencil(x1, x2, x3, x4, x5, h1invsq, h2invsq):
    """``Lap2DStencil`` (src/PlainNodalLaplacian.jl:150-152)."""
    return (2 * h1invsq + 2 * h2invsq) * x1 - h1invsq * (x2 + x3) - h2invsq * (x4 + x5)


def mult_op_neumann(mesh: RegularMesh, x: np.ndarray) -> np.ndarray:
    """``multOpNeumann!`` (src/PlainNodalLaplacian.jl:155-188), 2-D only, ghost =
    self on every face (BC = 1).  Written as an explicit double loop over the
    nodes so that it is an independent statement of the BC=1 rows."""
    assert mesh.dim == 2
    n1, n2 = (int(v) for v in mesh.nodes)
    h1, h2 = 1.0 / mesh.h[0] ** 2, 1.0 / mesh.h[1] ** 2
    X = np.asarray(x).reshape((n1, n2), order="F")
    Y = np.zeros_like(X)
    for j in range(n2):
        for i in range(n1):
            xl = X[i - 1, j] if i > 0 else X[i, j]
            xr = X[i + 1, j] if i < n1 - 1 else X[i, j]
            xd = X[i, j - 1] if j > 0 else X[i, j]
            xu = X[i, j + 1] if j < n2 - 1 else X[i, j]
            Y[i, j] = lap2d_stencil(X[i, j], xl, xr, xd, xu, h1, h2)
    return Y.ravel(order="F")


# --------------------------------------------------------------------------
# src/GetHelmholtz.jl
# --------------------------------------------------------------------------
def get_abl(n_nodes, neumann_at_first_dim: bool, abl_pad, abl_amp: float, code=None) -> np.ndarray:
    """``getABL`` (src/GetHelmholtz.jl:97-220), ``impl = 1`` for 2-D
    (:141-163) and the 3-D branch (:164-218).  Returns the nodal array shaped
    ``tuple(n_nodes)`` (column-major semantics: axis 0 = x)."""
    n = [int(v) for v in np.asarray(n_nodes).ravel()]
    pad = [int(v) for v in np.asarray(abl_pad).ravel()]
    if code is None:
        code = np.ones((len(n), 2), dtype=bool)
    else:
        code = np.array(code, dtype=bool)
    if len(n) == 2:
        n1, n2 = n
        p1, p2 = pad[0], pad[1]
        g = np.zeros((n1, n2))
        b_bwd1 = np.arange(p1, 0, -1, dtype=float) ** 2 / p1**2
        b_bwd2 = np.arange(p2, 0, -1, dtype=float) ** 2 / p2**2
        b_fwd1 = np.arange(1, p1 + 1, dtype=float) ** 2 / p1**2
        b_fwd2 = np.arange(1, p2 + 1, dtype=float) ** 2 / p2**2
        I1 = slice(n1 - p1, n1)
        I2 = slice(n2 - p2, n2)
        if not neumann_at_first_dim:
            g[:, :p2] += np.outer(np.ones(n1), b_bwd2)
            g[:p1, :p2] -= np.outer(b_bwd1, b_bwd2)
            g[I1, :p2] -= np.outer(b_fwd1, b_bwd2)
        g[:, I2] += np.outer(np.ones(n1), b_fwd2)
        g[:p1, :] += np.outer(b_bwd1, np.ones(n2))
        g[I1, :] += np.outer(b_fwd1, np.ones(n2))
        g[:p1, I2] -= np.outer(b_bwd1, b_fwd2)
        g[I1, I2] -= np.outer(b_fwd1, b_fwd2)
        return g * abl_amp
    # 3-D
    n1, n2, n3 = n
    x1 = np.linspace(-1.0, 1.0, n1)
    x2 = np.linspace(-1.0, 1.0, n2)
    x3 = np.linspace(0.0, 1.0, n3)
    X1, X2, X3 = np.meshgrid(x1, x2, x3, indexing="ij")

    def axis_profile(X, xv, p, axis, lo_on, hi_on):
        g = np.zeros(X.shape)
        if lo_on:
            gl = (X - xv[p - 1]) ** 2
            idx = [slice(None)] * 3
            idx[axis] = slice(p, None)
            gl[tuple(idx)] = 0.0
            g += gl
        if hi_on:
            gr = (X - xv[len(xv) - p]) ** 2
            idx = [slice(None)] * 3
            idx[axis] = slice(0, len(xv) - p)
            gr[tuple(idx)] = 0.0
            g += gr
        return g / (g.max() + 1e-5)

    if neumann_at_first_dim:
        code[2, 0] = False
    gx = axis_profile(X1, x1, pad[0], 0, code[0, 0], code[0, 1])
    gy = axis_profile(X2, x2, pad[1], 1, code[1, 0], code[1, 1])
    gz = axis_profile(X3, x3, pad[2], 2, code[2, 0], code[2, 1])
    g = (gx + gy + gz) * abl_amp
    g[g >= abl_amp] = abl_amp
    return g


def get_sommerfeld_bc(mesh: RegularMesh, m_nodal, omega: float, neumann_on_top: bool,
                      order_neumann_bc: int = 2) -> np.ndarray:
    """``getSommerfeldBC`` (src/GetHelmholtz.jl:222-247).  Returns the nodal
    complex array S; the operator subtracts it from the mass diagonal."""
    bc = get_bc(order_neumann_bc)
    shape = tuple(int(v) for v in mesh.nodes)
    M = np.asarray(m_nodal, dtype=float).reshape(shape, order="F")
    S = np.zeros(shape, dtype=complex)
    h = mesh.h
    c = lambda d: -1j * omega * (bc / h[d])
    if mesh.dim == 2:
        if not neumann_on_top:
            S[:, 0] += c(1) * np.sqrt(M[:, 0])
        S[:, -1] += c(1) * np.sqrt(M[:, -1])
        S[-1, :] += c(0) * np.sqrt(M[-1, :])
        S[0, :] += c(0) * np.sqrt(M[0, :])
    else:
        if not neumann_on_top:
            S[:, :, 0] += c(2) * np.sqrt(M[:, :, 0])
        S[:, :, -1] += c(2) * np.sqrt(M[:, :, -1])
        S[:, 0, :] += c(1) * np.sqrt(M[:, 0, :])
        S[:, -1, :] += c(1) * np.sqrt(M[:, -1, :])
        S[0, :, :] += c(0) * np.sqrt(M[0, :, :])
        S[-1, :, :] += c(0) * np.sqrt(M[-1, :, :])
    return S


def helmholtz_mass(mesh, m_nodal, omega, gamma, neumann_on_top, sommerfeld) -> np.ndarray:
    """The diagonal term of ``GetHelmholtzOperator`` (src/GetHelmholtz.jl:41-47):
    -omega^2 m (1 - i gamma / Re omega) - Sommerfeld."""
    m = np.asarray(m_nodal, dtype=float).ravel(order="F")
    g = np.asarray(gamma, dtype=float).ravel(order="F")
    w = complex(omega)
    mass = -(w**2) * m * (1.0 - 1j * g / w.real)
    if sommerfeld:
        mass = mass - get_sommerfeld_bc(mesh, m, w.real, neumann_on_top).ravel(order="F")
    return mass


def get_helmholtz_operator7(mesh, m_nodal, omega, gamma, neumann_at_first_dim: bool,
                            sommerfeld: bool, order_neumann_bc: int = 2) -> sp.csr_matrix:
    """7-argument ``GetHelmholtzOperator`` (src/GetHelmholtz.jl:33-50):
    H = Lap + diag(mass)."""
    L = get_nodal_laplacian_matrix(mesh, order_neumann_bc)
    mass = helmholtz_mass(mesh, m_nodal, omega, gamma, neumann_at_first_dim, sommerfeld)
    return (L.astype(complex) + sp.diags(mass)).tocsr()


def get_helmholtz_operator(mesh, m_nodal, omega, gamma, neumann_at_first_dim: bool, abl_pad,
                           abl_amp: float, sommerfeld: bool, order_neumann_bc: int = 2):
    """8-argument ``GetHelmholtzOperator`` (src/GetHelmholtz.jl:22-31): adds the
    ABL to ``gamma`` (or uses it alone when gamma is empty); returns (H, gamma)."""
    abl = get_abl(mesh.nodes, neumann_at_first_dim, abl_pad, abl_amp)
    if gamma is None or np.size(gamma) == 0:
        g = abl
    else:
        g = np.asarray(gamma, dtype=float).reshape(abl.shape, order="F") + abl
    H = get_helmholtz_operator7(mesh, m_nodal, omega, g, neumann_at_first_dim, sommerfeld,
                                order_neumann_bc)
    return H, g


def get_helmholtz_shift_op(m_nodal, omega: float, shift: float) -> sp.csr_matrix:
    """``GetHelmholtzShiftOP`` (src/GetHelmholtz.jl:81-83): diag(i shift omega^2 m)."""
    m = np.asarray(m_nodal, dtype=float).ravel(order="F")
    return sp.diags(m * (1j * shift * omega**2)).tocsr()


def get_maximal_frequency(m, mesh: RegularMesh) -> float:
    """``getMaximalFrequency`` (src/GetHelmholtz.jl:75-79): 10 points per
    wavelength at the slowest velocity."""
    return (0.1 * 2 * math.pi) / (float(np.max(mesh.h)) * math.sqrt(float(np.max(m))))


def helmholtz_fun(SHt: sp.spmatrix, shift_mat: sp.spmatrix):
    """``getHelmholtzFun`` (src/GetHelmholtz.jl:85-95) with the SpMatMul
    semantics y <- A^H x (two products: SH x then the shift op's adjoint)."""
    SH = SHt.conj().T.tocsr()
    Sh = shift_mat.conj().T.tocsr()
    return lambda x: SH @ x + Sh @ x


# --------------------------------------------------------------------------
# src/Helmholtz.jl:13-34
# --------------------------------------------------------------------------
@dataclass
class HelmholtzParam:
    Mesh: RegularMesh
    gamma: np.ndarray
    m: np.ndarray
    omega: complex
    NeumannOnTop: bool
    Sommerfeld: bool


def get_shifted_helmholtz_param(p: HelmholtzParam, s: float) -> HelmholtzParam:
    """``getShiftedHelmholtzParam`` (src/Helmholtz.jl:32-34): gamma + s Re(omega)."""
    return HelmholtzParam(p.Mesh, np.asarray(p.gamma, dtype=float) + s * complex(p.omega).real,
                          p.m, p.omega, p.NeumannOnTop, p.Sommerfeld)


def operator_from_param(p: HelmholtzParam, order_neumann_bc: int = 2) -> sp.csr_matrix:
    """``GetHelmholtzOperator(Hparam)`` (src/GetHelmholtz.jl:14-16)."""
    return get_helmholtz_operator7(p.Mesh, p.m, p.omega, p.gamma, p.NeumannOnTop, p.Sommerfeld,
                                   order_neumann_bc)


# --------------------------------------------------------------------------
# src/getPointSource.jl:63-111
# --------------------------------------------------------------------------
def get_top_point_src(mesh: RegularMesh):
    """``getTopPointSrc`` (src/getPointSource.jl:63-70), 1-based subscripts."""
    if mesh.dim == 3:
        return [(mesh.n[0] + 1) // 2, (mesh.n[1] + 1) // 2, 1]
    return [(mesh.n[0] + 1) // 2, 1]


def get_mid_point_src(mesh: RegularMesh):
    """``getMidPointSrc`` (src/getPointSource.jl:72-79), 1-based subscripts."""
    return [int(v + 1) // 2 for v in mesh.n]


def loc2cs(n_nodes, sub) -> int:
    """``loc2cs``/``loc2cs3D`` (src/getPointSource.jl:82-100): 1-based linear index."""
    n = [int(v) for v in n_nodes]
    if len(sub) == 2:
        return int(sub[0] + (sub[1] - 1) * n[0])
    return int(sub[0] + (sub[1] - 1) * n[0] + (sub[2] - 1) * n[0] * n[1])


def get_acoustic_point_source(mesh: RegularMesh, src=None, dtype=complex):
    """``getAcousticPointSource`` (src/getPointSource.jl:105-111): q = e_src / ||h||^2."""
    if src is None:
        src = get_top_point_src(mesh)
    nn = [int(v) for v in mesh.nodes]
    q = np.zeros(int(np.prod(nn)), dtype=dtype)
    q[loc2cs(nn, src) - 1] = 1.0 / float(np.linalg.norm(mesh.h)) ** 2
    return q, src
