"""Oracle: reference elements, Lagrange shape functions and Gauss rules.

Restates (TEST INFRASTRUCTURE ONLY, see oracle/__init__.py):
  * reference_shape_value for Lagrange{RefLine,1|2}, {RefQuadrilateral,1|2},
    {RefHexahedron,1|2}: src/interpolations.jl:600-659, :1038-1159
  * reference_shape_gradient by forward-mode dual numbers, as Tensors.jl's
    `gradient` does it: src/interpolations.jl:280-302
  * vectorized interpolations ip^vdim: src/interpolations.jl:1769-1825
  * entity dof tables (vertex/edge/face/volume interior indices):
    src/interpolations.jl:584-590, :619, :666-667, :1075-1085
  * reference entities of cells: src/Grid/grid.jl:194-229, :261-275
  * Gauss-Legendre by Golub-Welsch (implicit QL, EISPACK imtql2 variant):
    src/Quadrature/generate_quadrature.jl:93-325
  * tensor-product QuadratureRule on hypercubes: src/Quadrature/quadrature.jl:83-107
  * FacetQuadratureRule + facet_to_element_transformation:
    src/Quadrature/quadrature.jl:223-227, src/FEValues/facet_integrals.jl:103-202
"""
import math

import numpy as np


# ----------------------------------------------------------------------------
# Forward-mode dual numbers (value + gradient), the arithmetic Tensors.jl uses
# for `gradient(x -> reference_shape_value(ip, x, i), ξ)`.
# ----------------------------------------------------------------------------
class Dual:
    __slots__ = ("v", "g")

    def __init__(self, v, g):
        self.v = float(v)
        self.g = tuple(float(x) for x in g)

    @staticmethod
    def _lift(o, n):
        return o if isinstance(o, Dual) else Dual(o, (0.0,) * n)

    def __add__(self, o):
        o = Dual._lift(o, len(self.g))
        return Dual(self.v + o.v, tuple(a + b for a, b in zip(self.g, o.g)))

    __radd__ = __add__

    def __sub__(self, o):
        o = Dual._lift(o, len(self.g))
        return Dual(self.v - o.v, tuple(a - b for a, b in zip(self.g, o.g)))

    def __rsub__(self, o):
        o = Dual._lift(o, len(self.g))
        return o - self

    def __neg__(self):
        return Dual(-self.v, tuple(-a for a in self.g))

    def __mul__(self, o):
        if not isinstance(o, Dual):
            return Dual(self.v * o, tuple(a * o for a in self.g))
        return Dual(self.v * o.v, tuple(a * o.v + self.v * b for a, b in zip(self.g, o.g)))

    __rmul__ = __mul__

    def __truediv__(self, o):
        assert not isinstance(o, Dual)
        return Dual(self.v / o, tuple(a / o for a in self.g))

    def __pow__(self, n):
        assert n == 2
        return self * self


# ----------------------------------------------------------------------------
# Reference shapes and cell entities (grid.jl:194-229)
# ----------------------------------------------------------------------------
REF_VERTICES = {
    "RefLine": 2, "RefQuadrilateral": 4, "RefHexahedron": 8,
}
REF_EDGES = {
    "RefLine": ((1, 2),),
    "RefQuadrilateral": ((1, 2), (2, 3), (3, 4), (4, 1)),
    "RefHexahedron": ((1, 2), (2, 3), (3, 4), (4, 1), (5, 6), (6, 7),
                      (7, 8), (8, 5), (1, 5), (2, 6), (3, 7), (4, 8)),
}
REF_FACES = {
    "RefLine": (),
    "RefQuadrilateral": ((1, 2, 3, 4),),
    "RefHexahedron": ((1, 4, 3, 2), (1, 2, 6, 5), (2, 3, 7, 6),
                      (3, 4, 8, 7), (1, 5, 8, 4), (5, 6, 7, 8)),
}
REF_DIM = {"RefLine": 1, "RefQuadrilateral": 2, "RefHexahedron": 3}


def reference_face_edgenrs(shape):
    """grid.jl:261-275: for each face, the edge number of each consecutive vertex pair."""
    out = []
    edges = REF_EDGES[shape]
    for face in REF_FACES[shape]:
        nrs = []
        for j in range(len(face)):
            v1, v2 = face[j], face[(j + 1) % len(face)]
            for k, e in enumerate(edges):
                if e == (v1, v2) or e == (v2, v1):
                    nrs.append(k + 1)
                    break
        out.append(tuple(nrs))
    return tuple(out)


def reference_facets(shape):
    """Facets are edges in 2D, faces in 3D, vertices in 1D."""
    d = REF_DIM[shape]
    if d == 3:
        return REF_FACES[shape]
    if d == 2:
        return REF_EDGES[shape]
    return ((1,), (2,))


# ----------------------------------------------------------------------------
# Lagrange interpolations
# ----------------------------------------------------------------------------
class Lagrange:
    """Scalar Lagrange{shape, order} (order 1 or 2) on line/quad/hex."""

    def __init__(self, shape, order):
        assert shape in REF_DIM and order in (1, 2)
        self.shape = shape
        self.order = order
        self.rdim = REF_DIM[shape]

    def __eq__(self, o):
        return isinstance(o, Lagrange) and (o.shape, o.order) == (self.shape, self.order)

    def __hash__(self):
        return hash((self.shape, self.order))

    def __pow__(self, vdim):
        return VectorizedInterpolation(self, vdim)

    @property
    def nbasefunctions(self):
        return {("RefLine", 1): 2, ("RefLine", 2): 3, ("RefQuadrilateral", 1): 4,
                ("RefQuadrilateral", 2): 9, ("RefHexahedron", 1): 8,
                ("RefHexahedron", 2): 27}[(self.shape, self.order)]

    # interpolations.jl:584-590, 619, 666-667, 1075-1085
    def vertexdof_indices(self):
        return tuple((i,) for i in range(1, REF_VERTICES[self.shape] + 1))

    def edgedof_interior_indices(self):
        ne = len(REF_EDGES[self.shape])
        if self.order == 1:
            return tuple(() for _ in range(ne))
        if self.shape == "RefLine":
            return ((3,),)
        if self.shape == "RefQuadrilateral":
            return ((5,), (6,), (7,), (8,))
        return tuple((i,) for i in range(9, 21))

    def facedof_interior_indices(self):
        nf = len(REF_FACES[self.shape])
        if self.order == 1:
            return tuple(() for _ in range(nf))
        if self.shape == "RefQuadrilateral":
            return ((9,),)
        if self.shape == "RefHexahedron":
            return tuple((i,) for i in range(21, 27))
        return ()

    def volumedof_interior_indices(self):
        if self.order == 2 and self.shape == "RefHexahedron":
            return (27,)
        return ()

    # interpolations.jl:355-427 (edgedof_indices / facedof_indices)
    def edgedof_indices(self):
        vd, ed = self.vertexdof_indices(), self.edgedof_interior_indices()
        out = []
        for k, e in enumerate(REF_EDGES[self.shape]):
            t = []
            for v in e:
                t.extend(vd[v - 1])
            t.extend(ed[k])
            out.append(tuple(t))
        return tuple(out)

    def facedof_indices(self):
        vd, ed, fd = (self.vertexdof_indices(), self.edgedof_interior_indices(),
                      self.facedof_interior_indices())
        fe = reference_face_edgenrs(self.shape)
        out = []
        for k, f in enumerate(REF_FACES[self.shape]):
            t = []
            for v in f:
                t.extend(vd[v - 1])
            for en in fe[k]:
                t.extend(ed[en - 1])
            t.extend(fd[k])
            out.append(tuple(t))
        return tuple(out)

    def dirichlet_facetdof_indices(self):
        """interpolations.jl:474-482."""
        if self.rdim == 3:
            return self.facedof_indices()
        if self.rdim == 2:
            return self.edgedof_indices()
        return self.vertexdof_indices()

    # interpolations.jl:640-659, 1040-1066, 1073-1159 (and Line 600-625)
    def reference_coordinates(self):
        if self.shape == "RefLine":
            c = [(-1.0,), (1.0,)]
            return c + [(0.0,)] if self.order == 2 else c
        if self.shape == "RefQuadrilateral":
            c = [(-1.0, -1.0), (1.0, -1.0), (1.0, 1.0), (-1.0, 1.0)]
            if self.order == 2:
                c += [(0.0, -1.0), (1.0, 0.0), (0.0, 1.0), (-1.0, 0.0), (0.0, 0.0)]
            return c
        c = [(-1.0, -1.0, -1.0), (1.0, -1.0, -1.0), (1.0, 1.0, -1.0), (-1.0, 1.0, -1.0),
             (-1.0, -1.0, 1.0), (1.0, -1.0, 1.0), (1.0, 1.0, 1.0), (-1.0, 1.0, 1.0)]
        if self.order == 2:
            c += [(0.0, -1.0, -1.0), (1.0, 0.0, -1.0), (0.0, 1.0, -1.0), (-1.0, 0.0, -1.0),
                  (0.0, -1.0, 1.0), (1.0, 0.0, 1.0), (0.0, 1.0, 1.0), (-1.0, 0.0, 1.0),
                  (-1.0, -1.0, 0.0), (1.0, -1.0, 0.0), (1.0, 1.0, 0.0), (-1.0, 1.0, 0.0),
                  (0.0, 0.0, -1.0), (0.0, -1.0, 0.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0),
                  (-1.0, 0.0, 0.0), (0.0, 0.0, 1.0), (0.0, 0.0, 0.0)]
        return c

    def shape_value(self, xi, i):
        """reference_shape_value(ip, ξ, i), 1-based i; xi entries may be Dual."""
        s, o = self.shape, self.order
        if s == "RefLine":
            x = xi[0]
            if o == 1:
                return (1 - x) / 2 if i == 1 else (1 + x) / 2
            return (x * (x - 1) / 2, x * (x + 1) / 2, 1 - x ** 2)[i - 1]
        if s == "RefQuadrilateral":
            x, y = xi
            if o == 1:
                return ((1 - x) * (1 - y) / 4, (1 + x) * (1 - y) / 4,
                        (1 + x) * (1 + y) / 4, (1 - x) * (1 + y) / 4)[i - 1]
            return _q2_quad(x, y, i)
        x, y, z = xi
        if o == 1:
            return ((1 - x) * (1 - y) * (1 - z) / 8, (1 + x) * (1 - y) * (1 - z) / 8,
                    (1 + x) * (1 + y) * (1 - z) / 8, (1 - x) * (1 + y) * (1 - z) / 8,
                    (1 - x) * (1 - y) * (1 + z) / 8, (1 + x) * (1 - y) * (1 + z) / 8,
                    (1 + x) * (1 + y) * (1 + z) / 8, (1 - x) * (1 + y) * (1 + z) / 8)[i - 1]
        return _q2_hex(x, y, z, i)

    def shape_gradient_and_value(self, xi, i):
        n = len(xi)
        duals = [Dual(xi[d], tuple(1.0 if e == d else 0.0 for e in range(n))) for d in range(n)]
        r = self.shape_value(duals, i)
        if not isinstance(r, Dual):
            return np.zeros(n), float(r)
        return np.array(r.g), r.v


def _q2_quad(x, y, i):
    # interpolations.jl:686-698 (Lagrange{RefQuadrilateral,2})
    return ((x ** 2 - x) * (y ** 2 - y) / 4, (x ** 2 + x) * (y ** 2 - y) / 4,
            (x ** 2 + x) * (y ** 2 + y) / 4, (x ** 2 - x) * (y ** 2 + y) / 4,
            (1 - x ** 2) * (y ** 2 - y) / 2, (x ** 2 + x) * (1 - y ** 2) / 2,
            (1 - x ** 2) * (y ** 2 + y) / 2, (x ** 2 - x) * (1 - y ** 2) / 2,
            (1 - x ** 2) * (1 - y ** 2))[i - 1]


def _q2_hex(x, y, z, i):
    # interpolations.jl:1110-1159
    def f1(t):
        return -t * (1 - t) / 2

    def f2(t):
        return (1 + t) * (1 - t)

    def f3(t):
        return t * (1 + t) / 2
    table = {
        1: (f1, f1, f1), 2: (f3, f1, f1), 3: (f3, f3, f1), 4: (f1, f3, f1),
        5: (f1, f1, f3), 6: (f3, f1, f3), 7: (f3, f3, f3), 8: (f1, f3, f3),
        9: (f2, f1, f1), 10: (f3, f2, f1), 11: (f2, f3, f1), 12: (f1, f2, f1),
        13: (f2, f1, f3), 14: (f3, f2, f3), 15: (f2, f3, f3), 16: (f1, f2, f3),
        17: (f1, f1, f2), 18: (f3, f1, f2), 19: (f3, f3, f2), 20: (f1, f3, f2),
        21: (f2, f2, f1), 22: (f2, f1, f2), 23: (f3, f2, f2), 24: (f2, f3, f2),
        25: (f1, f2, f2), 26: (f2, f2, f3), 27: (f2, f2, f2),
    }
    a, b, c = table[i]
    return a(x) * b(y) * c(z)


class VectorizedInterpolation:
    """ip^vdim (interpolations.jl:1769-1825): basis I -> node (I-1)//vdim+1, comp (I-1)%vdim+1."""

    def __init__(self, ip, vdim):
        self.ip = ip
        self.vdim = vdim
        self.shape = ip.shape
        self.rdim = ip.rdim
        self.order = ip.order

    def __eq__(self, o):
        return isinstance(o, VectorizedInterpolation) and o.ip == self.ip and o.vdim == self.vdim

    def __hash__(self):
        return hash((self.ip, self.vdim))

    @property
    def nbasefunctions(self):
        return self.vdim * self.ip.nbasefunctions


def base_interpolation(ip):
    return ip.ip if isinstance(ip, VectorizedInterpolation) else ip


def n_copies(ip):
    return ip.vdim if isinstance(ip, VectorizedInterpolation) else 1


def geometric_interpolation(shape):
    return Lagrange(shape, 1)


# ----------------------------------------------------------------------------
# Quadrature: Golub-Welsch (generate_quadrature.jl:93-325)
# ----------------------------------------------------------------------------
def _steig(d, e, z, maxits=50):
    """generate_quadrature.jl:227-305 (implicit QL, first eigenvector components)."""
    n = len(z)
    z[0] = 1.0
    for i in range(1, n):
        z[i] = 0.0
    e[n - 1] = 0.0
    if n == 1:
        return
    eps = np.finfo(np.float64).eps
    for l in range(n):
        for j in range(1, maxits + 1):
            m = n - 1
            for i in range(l, n - 1):
                if abs(e[i]) <= eps * (abs(d[i]) + abs(d[i + 1])):
                    m = i
                    break
            p = d[l]
            if m == l:
                continue
            if j == maxits:
                raise RuntimeError("No convergence")
            g = (d[l + 1] - p) / (2 * e[l])
            r = math.hypot(g, 1.0)
            g = d[m] - p + e[l] / (g + math.copysign(r, g))
            s = 1.0
            c = 1.0
            p = 0.0
            for i in range(m - 1, l - 1, -1):
                f = s * e[i]
                b = c * e[i]
                if abs(f) < abs(g):
                    s = f / g
                    r = math.hypot(s, 1.0)
                    e[i + 1] = g * r
                    c = 1.0 / r
                    s *= c
                else:
                    c = g / f
                    r = math.hypot(c, 1.0)
                    e[i + 1] = f * r
                    s = 1.0 / r
                    c *= s
                g = d[i + 1] - p
                r = (d[i] - g) * s + 2 * c * b
                p = s * r
                d[i + 1] = g + p
                g = c * r - b
                f = z[i + 1]
                z[i + 1] = s * z[i] + c * f
                z[i] = c * z[i] - s * f
            d[l] -= p
            e[l] = g
            e[m] = 0.0


def gauss_legendre(n):
    """GaussQuadrature.legendre(Float64, n) (generate_quadrature.jl:93-113, 150-199)."""
    a = [0.0] * n
    b = [i / math.sqrt(float(4 * i * i - 1)) for i in range(1, n + 1)]
    w = [0.0] * n
    _steig(a, b, w)
    w = [2.0 * wi * wi for wi in w]
    idx = sorted(range(n), key=lambda k: a[k])
    return [a[k] for k in idx], [w[k] for k in idx]


class QuadratureRule:
    """QuadratureRule{RefHypercube{dim}}(:legendre, order) (quadrature.jl:83-107).

    points[q] are tuples (first coordinate fastest); weight = 1.0*w[i1]*w[i2]*...
    """

    def __init__(self, shape, order):
        self.shape = shape
        dim = REF_DIM[shape]
        p, w = gauss_legendre(order)
        self.points = []
        self.weights = []
        for idx in _product_first_fastest([order] * dim):
            self.points.append(tuple(p[i] for i in idx))
            weight = 1.0
            for i in idx:
                weight *= w[i]
            self.weights.append(weight)

    @property
    def nqp(self):
        return len(self.weights)


def _product_first_fastest(dims):
    """Iterators.product order: first index fastest (0-based tuples)."""
    total = 1
    for d in dims:
        total *= d
    for lin in range(total):
        idx = []
        r = lin
        for d in dims:
            idx.append(r % d)
            r //= d
        yield tuple(idx)


def facet_to_element(point, shape, facet):
    """facet_integrals.jl:71-177 (1-based facet)."""
    if shape == "RefQuadrilateral":
        x = point[0]
        return ((x, -1.0), (1.0, x), (-x, 1.0), (-1.0, -x))[facet - 1]
    if shape == "RefHexahedron":
        x, y = point
        return ((y, x, -1.0), (x, -1.0, y), (1.0, x, y), (-x, 1.0, y),
                (-1.0, y, x), (x, y, 1.0))[facet - 1]
    raise ValueError(shape)


def weighted_normal(J, shape, facet):
    """facet_integrals.jl:122-131, 191-202; J is an (sdim, rdim) numpy array."""
    if shape == "RefQuadrilateral":
        return np.array(((J[1, 0], -J[0, 0]), (J[1, 1], -J[0, 1]),
                         (-J[1, 0], J[0, 0]), (-J[1, 1], J[0, 1]))[facet - 1])
    pairs = ((1, 0), (0, 2), (1, 2), (2, 0), (2, 1), (0, 1))
    a, b = pairs[facet - 1]
    return cross(J[:, a], J[:, b])


def cross(u, v):
    return np.array([u[1] * v[2] - u[2] * v[1],
                     u[2] * v[0] - u[0] * v[2],
                     u[0] * v[1] - u[1] * v[0]])


class FacetQuadratureRule:
    """FacetQuadratureRule{shape}(order): the facet-shape Gauss rule mapped onto every facet."""

    def __init__(self, shape, order):
        self.shape = shape
        fshape = {"RefQuadrilateral": "RefLine", "RefHexahedron": "RefQuadrilateral"}[shape]
        qr = QuadratureRule(fshape, order)
        nf = len(reference_facets(shape))
        self.rules = []
        for f in range(1, nf + 1):
            pts = [facet_to_element(p, shape, f) for p in qr.points]
            self.rules.append((pts, list(qr.weights)))
