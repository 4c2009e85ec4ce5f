"""Reference shapes, Lagrange interpolations and Gauss rules of the host interface.

Mirrors Ferrite's `Lagrange{RefShape, order}` (src/interpolations.jl:584-667,
:1038-1159), vectorized `ip^vdim` (:1769-1825), `QuadratureRule{RefShape}(order)`
(src/Quadrature/quadrature.jl:83-107) and `FacetQuadratureRule` (:223-227).

Shape functions are evaluated as tensor products of 1D Lagrange polynomials on
{-1, 1} (order 1) or {-1, 0, 1} (order 2) at each basis function's reference
coordinate; gradients are the analytic derivatives.  Gauss points come from
the Golub–Welsch eigenproblem of the Legendre Jacobi matrix (as the reference
does, here with LAPACK's symmetric tridiagonal solver).
"""
import numpy as np

RefLine, RefQuadrilateral, RefHexahedron = "RefLine", "RefQuadrilateral", "RefHexahedron"
REF_DIM = {RefLine: 1, RefQuadrilateral: 2, RefHexahedron: 3}

# Node ordering tables: basis function i sits at REF_COORDS[(shape, order)][i]
# (interpolations.jl:600-620, 640-680, 1040-1107): vertices, edges, faces, volume.
_Q1 = {
    RefLine: [(-1,), (1,)],
    RefQuadrilateral: [(-1, -1), (1, -1), (1, 1), (-1, 1)],
    RefHexahedron: [(-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1),
                    (-1, -1, 1), (1, -1, 1), (1, 1, 1), (-1, 1, 1)],
}
REF_COORDS = {
    (RefLine, 1): _Q1[RefLine],
    (RefLine, 2): _Q1[RefLine] + [(0,)],
    (RefQuadrilateral, 1): _Q1[RefQuadrilateral],
    (RefQuadrilateral, 2): _Q1[RefQuadrilateral] + [(0, -1), (1, 0), (0, 1), (-1, 0), (0, 0)],
    (RefHexahedron, 1): _Q1[RefHexahedron],
    (RefHexahedron, 2): _Q1[RefHexahedron] + [
        (0, -1, -1), (1, 0, -1), (0, 1, -1), (-1, 0, -1), (0, -1, 1), (1, 0, 1), (0, 1, 1), (-1, 0, 1),
        (-1, -1, 0), (1, -1, 0), (1, 1, 0), (-1, 1, 0),
        (0, 0, -1), (0, -1, 0), (1, 0, 0), (0, 1, 0), (-1, 0, 0), (0, 0, 1),
        (0, 0, 0)],
}

# Reference facets (grid.jl:194-229): local vertex numbers (1-based) of each facet.
REF_FACETS = {
    RefQuadrilateral: ((1, 2), (2, 3), (3, 4), (4, 1)),
    RefHexahedron: ((1, 4, 3, 2), (1, 2, 6, 5), (2, 3, 7, 6), (3, 4, 8, 7), (1, 5, 8, 4), (5, 6, 7, 8)),
}


def _lag1d(order, node, t):
    """1D Lagrange basis on {-1,1} / {-1,0,1} attached to `node`, value and derivative at t."""
    t = np.asarray(t, dtype=np.float64)
    if order == 1:
        s = float(node)
        return (1 + s * t) / 2, np.full_like(t, s / 2)
    if node == -1:
        return t * (t - 1) / 2, t - 0.5
    if node == 1:
        return t * (t + 1) / 2, t + 0.5
    return 1 - t * t, -2 * t


class Lagrange:
    """Scalar Lagrange{shape, order}, order 1 or 2."""

    def __init__(self, shape, order=1):
        if shape not in REF_DIM or order not in (1, 2):
            raise ValueError(f"unsupported interpolation Lagrange{{{shape},{order}}}")
        self.shape, self.order, self.rdim = shape, order, REF_DIM[shape]

    def __repr__(self):
        return f"Lagrange{{{self.shape}, {self.order}}}()"

    def __eq__(self, o):
        return isinstance(o, Lagrange) and (o.shape, o.order) == (self.shape, self.order)

    def __hash__(self):
        return hash((self.shape, self.order))

    def __pow__(self, vdim):
        return VectorizedInterpolation(self, vdim)

    @property
    def nbasefunctions(self):
        return len(REF_COORDS[(self.shape, self.order)])

    def reference_coordinates(self):
        return np.array(REF_COORDS[(self.shape, self.order)], dtype=np.float64)

    def values_and_gradients(self, points):
        """N[q, i] and dNdξ[q, i, d] at the given reference points (nq, rdim)."""
        pts = np.atleast_2d(np.asarray(points, dtype=np.float64))
        nodes = self.reference_coordinates()
        nq, nb, rd = pts.shape[0], nodes.shape[0], self.rdim
        N = np.ones((nq, nb))
        dN = np.ones((nq, nb, rd))
        for i in range(nb):
            vals, ders = [], []
            for d in range(rd):
                v, g = _lag1d(self.order, nodes[i, d], pts[:, d])
                vals.append(v)
                ders.append(g)
            for d in range(rd):
                N[:, i] *= vals[d]
                for e in range(rd):
                    dN[:, i, e] *= ders[d] if d == e else vals[d]
        return N, dN

    def facet_dof_indices(self):
        """dirichlet_facetdof_indices (interpolations.jl:474-482): 1-based local basis
        functions on each facet: facet vertices, then facet edges, then facet interior."""
        if self.shape == RefQuadrilateral:
            edge_mid = {(1, 2): 5, (2, 3): 6, (3, 4): 7, (4, 1): 8}
            out = []
            for f in REF_FACETS[RefQuadrilateral]:
                t = list(f)
                if self.order == 2:
                    t.append(edge_mid[f])
                out.append(tuple(t))
            return tuple(out)
        if self.shape == RefHexahedron:
            edges = ((1, 2), (2, 3), (3, 4), (4, 1), (5, 6), (6, 7), (7, 8), (8, 5), (1, 5), (2, 6), (3, 7), (4, 8))
            out = []
            for k, f in enumerate(REF_FACETS[RefHexahedron]):
                t = list(f)
                if self.order == 2:
                    for j in range(4):
                        a, b = f[j], f[(j + 1) % 4]
                        e = next(i for i, ed in enumerate(edges) if ed in ((a, b), (b, a)))
                        t.append(9 + e)
                    t.append(21 + k)
                out.append(tuple(t))
            return tuple(out)
        return ((1,), (2,))


class VectorizedInterpolation:
    """ip^vdim: basis I (1-based) -> node (I-1)//vdim + 1, component (I-1) % vdim + 1."""

    def __init__(self, ip, vdim):
        self.ip, self.vdim = ip, vdim
        self.shape, self.order, self.rdim = ip.shape, ip.order, ip.rdim

    def __repr__(self):
        return f"{self.ip!r}^{self.vdim}"

    def __eq__(self, o):
        return isinstance(o, VectorizedInterpolation) and (o.ip, o.vdim) == (self.ip, self.vdim)

    def __hash__(self):
        return hash((self.ip, self.vdim))

    @property
    def nbasefunctions(self):
        return self.vdim * self.ip.nbasefunctions


def base_interpolation(ip):
    return ip.ip if isinstance(ip, VectorizedInterpolation) else ip


def n_components(ip):
    return ip.vdim if isinstance(ip, VectorizedInterpolation) else 1


def gauss_legendre(n):
    """n-point Gauss–Legendre rule on [-1, 1] by Golub–Welsch (points ascending)."""
    if n == 1:
        return np.array([0.0]), np.array([2.0])
    k = np.arange(1, n, dtype=np.float64)
    b = k / np.sqrt(4 * k * k - 1)
    T = np.diag(b, 1) + np.diag(b, -1)
    lam, V = np.linalg.eigh(T)
    w = 2.0 * V[0, :] ** 2
    idx = np.argsort(lam)
    return lam[idx], w[idx]


class QuadratureRule:
    """QuadratureRule{shape}(order): tensor Gauss rule, first coordinate fastest,
    weight = product of 1D weights (quadrature.jl:83-107)."""

    def __init__(self, shape, order):
        self.shape, self.order = shape, order
        dim = REF_DIM[shape]
        p, w = gauss_legendre(order)
        q = np.arange(order ** dim)
        idx = np.stack([(q // order ** d) % order for d in range(dim)], axis=1)  # first index fastest
        self.points = p[idx]
        weights = np.ones(idx.shape[0])
        for d in range(dim):
            weights = weights * w[idx[:, d]]
        self.weights = weights

    @property
    def nqp(self):
        return len(self.weights)


def facet_to_element(points2, shape, facet):
    """facet_to_element_transformation (facet_integrals.jl:103-177), 1-based facet."""
    p = np.atleast_2d(points2)
    if shape == RefQuadrilateral:
        x = p[:, 0]
        one = np.ones_like(x)
        return np.stack([(x, -one), (one, x), (-x, one), (-one, -x)][facet - 1], axis=1)
    x, y = p[:, 0], p[:, 1]
    one = np.ones_like(x)
    return np.stack([(y, x, -one), (x, -one, y), (one, x, y), (-x, one, y), (-one, y, x), (x, y, one)][facet - 1],
                    axis=1)


class FacetQuadratureRule:
    """FacetQuadratureRule{shape}(order): the facet Gauss rule mapped onto every facet."""

    def __init__(self, shape, order):
        self.shape, self.order = shape, order
        fshape = {RefQuadrilateral: RefLine, RefHexahedron: RefQuadrilateral}[shape]
        q = QuadratureRule(fshape, order)
        nf = len(REF_FACETS[shape])
        self.points = [facet_to_element(q.points, shape, f) for f in range(1, nf + 1)]
        self.weights = [q.weights.copy() for _ in range(nf)]
