"""CPU restatement of the strain / stress post-processing (numpy, FP64 with
the reference's FP32 accumulators).

TEST INFRASTRUCTURE ONLY: the checker for pdem_strain_stress.  The product
never imports it.

Follows, operation by operation:
  quadrature points      TriElem::init (src/fe/triElem.cpp:192-417),
                         TetElem::init (src/fe/tetElem.cpp:300-409)
  shape values           TriElem::getShapes (triElem.cpp:119-123),
                         TetElem::getShapes (tetElem.cpp:180-184)
  gradients, mapping     TriElem::getQuadDatas (triElem.cpp:58-90),
                         TetElem::getQuadDatas (tetElem.cpp:110-150) with
                         util::det / util::inv / util::dot (util/matrix.cpp:38-150)
  current points         fe::getCurrentQuadPoints (src/fe/meshUtil.cpp:176-278)
  strain, stress         fe::getStrainStress (meshUtil.cpp:280-420); strain and
                         stress are util::SymMatrix3, float storage
                         (util/matrix.h:257-300), order xx yy zz yz xz xy
"""
import numpy as np

F32 = np.float32


def quad_points(dim, order):
    """Reference-element quadrature points (3 coordinates each)."""
    if dim == 2:
        t = {
            1: [(1. / 3., 1. / 3.)],
            2: [(1. / 6., 1. / 6.), (2. / 3., 1. / 6.), (1. / 6., 2. / 3.)],
            3: [(1. / 3., 1. / 3.), (1. / 5., 3. / 5.), (1. / 5., 1. / 5.), (3. / 5., 1. / 5.)],
            4: [(0.44594849091597, 0.44594849091597), (0.44594849091597, 0.10810301816807),
                (0.10810301816807, 0.44594849091597), (0.09157621350977, 0.09157621350977),
                (0.09157621350977, 0.81684757298046), (0.81684757298046, 0.09157621350977)],
            5: [(0.33333333333333, 0.33333333333333), (0.47014206410511, 0.47014206410511),
                (0.47014206410511, 0.05971587178977), (0.05971587178977, 0.47014206410511),
                (0.10128650732346, 0.10128650732346), (0.10128650732346, 0.79742698535309),
                (0.79742698535309, 0.10128650732346)],
        }[order]
        return np.array([(a, b, 0.0) for a, b in t])
    a, b = 0.585410196624969, 0.138196601125011
    t = {
        1: [(1. / 4., 1. / 4., 1. / 4.)],
        2: [(a, b, b), (b, a, b), (b, b, a), (b, b, b)],
        3: [(0.25, 0.25, 0.25), (0.5, 1. / 6., 1. / 6.), (1. / 6., 0.5, 1. / 6.), (1. / 6., 1. / 6., 0.5),
            (1. / 6., 1. / 6., 1. / 6.)],
    }[order]
    return np.array(t)


def shapes(dim, p):
    if dim == 2:
        return np.array([1. - p[0] - p[1], p[0], p[1]])
    return np.array([1. - p[0] - p[1] - p[2], p[0], p[1], p[2]])


def gradients(dim, X):
    """dN [ne][nv][dim] of linear elements with vertex coordinates X [ne][nv][3]."""
    ne = X.shape[0]
    if dim == 2:
        J00 = X[:, 1, 0] - X[:, 0, 0]
        J01 = X[:, 1, 1] - X[:, 0, 1]
        J10 = X[:, 2, 0] - X[:, 0, 0]
        J11 = X[:, 2, 1] - X[:, 0, 1]
        det = J00 * J11 - J01 * J10
        gr = ((-1., -1.), (1., 0.), (0., 1.))
        dN = np.zeros((ne, 3, 2))
        for i, (g0, g1) in enumerate(gr):
            dN[:, i, 0] = (g0 * J11 - g1 * J01) / det
            dN[:, i, 1] = (-g0 * J10 + g1 * J00) / det
        return dN
    m = X[:, 1:, :] - X[:, :1, :]
    det = (m[:, 0, 0] * (m[:, 1, 1] * m[:, 2, 2] - m[:, 2, 1] * m[:, 1, 2])
           - m[:, 0, 1] * (m[:, 1, 0] * m[:, 2, 2] - m[:, 2, 0] * m[:, 1, 2])
           + m[:, 0, 2] * (m[:, 1, 0] * m[:, 2, 1] - m[:, 2, 0] * m[:, 1, 1]))
    di = 1. / det
    n = np.zeros((ne, 3, 3))
    n[:, 0, 0] = di * (m[:, 1, 1] * m[:, 2, 2] - m[:, 2, 1] * m[:, 1, 2])
    n[:, 0, 1] = -di * (m[:, 0, 1] * m[:, 2, 2] - m[:, 2, 1] * m[:, 0, 2])
    n[:, 0, 2] = di * (m[:, 0, 1] * m[:, 1, 2] - m[:, 1, 1] * m[:, 0, 2])
    n[:, 1, 0] = -di * (m[:, 1, 0] * m[:, 2, 2] - m[:, 2, 0] * m[:, 1, 2])
    n[:, 1, 1] = di * (m[:, 0, 0] * m[:, 2, 2] - m[:, 2, 0] * m[:, 0, 2])
    n[:, 1, 2] = -di * (m[:, 0, 0] * m[:, 1, 2] - m[:, 1, 0] * m[:, 0, 2])
    n[:, 2, 0] = di * (m[:, 1, 0] * m[:, 2, 1] - m[:, 2, 0] * m[:, 1, 1])
    n[:, 2, 1] = -di * (m[:, 0, 0] * m[:, 2, 1] - m[:, 2, 0] * m[:, 0, 1])
    n[:, 2, 2] = di * (m[:, 0, 0] * m[:, 1, 1] - m[:, 1, 0] * m[:, 0, 1])
    gr = ((-1., -1., -1.), (1., 0., 0.), (0., 1., 0.), (0., 0., 1.))
    dN = np.zeros((ne, 4, 3))
    for i, g in enumerate(gr):
        for k in range(3):
            r = np.zeros(ne)
            for j in range(3):
                r = r + n[:, k, j] * g[j]
            dN[:, i, k] = r
    return dN


def mapped_points(dim, X, N, qp):
    """Physical position of a reference point with shape values N."""
    if dim == 2:
        p = np.zeros((X.shape[0], 3))
        p[:, 0] = N[0] * X[:, 0, 0] + N[1] * X[:, 1, 0] + N[2] * X[:, 2, 0]
        p[:, 1] = N[0] * X[:, 0, 1] + N[1] * X[:, 1, 1] + N[2] * X[:, 2, 1]
        p[:, 2] = qp[2]
        return p
    p = np.zeros((X.shape[0], 3))
    for i in range(4):
        p = p + N[i] * X[:, i, :]
    return p


def strain_stress(dim, order, xr, u, cells, part, elastic):
    """xq [ne*nq][3], strain and stress [ne*nq][6] (float32).

    xr, u: [n][3] node fields (z = 0 in 2D); cells: [ne][dim+1] global node
    ids; part: [ne] particle ids; elastic: [np][4] nu, lambda, mu, plane strain."""
    cells = np.asarray(cells, np.int64)
    X = xr[cells]
    U = u[cells]
    ne, nv = cells.shape
    pts = quad_points(dim, order)
    nq = len(pts)
    dN = gradients(dim, X)
    em = np.asarray(elastic, np.float64)[np.asarray(part, np.int64)]
    nu, lam, mu, pstrain = em[:, 0], em[:, 1], em[:, 2], em[:, 3] != 0.
    xq = np.zeros((ne, nq, 3))
    strain = np.zeros((ne, nq, 6), F32)
    stress = np.zeros((ne, nq, 6), F32)

    def acc(a, t):
        return (a.astype(np.float64) + t).astype(F32)

    for q in range(nq):
        N = shapes(dim, pts[q])
        p = mapped_points(dim, X, N, pts[q])
        for i in range(nv):
            p = p + U[:, i, :] * N[i]
        xq[:, q] = p
        sn = [np.zeros(ne, F32) for _ in range(6)]
        for i in range(nv):
            ui = U[:, i, :]
            d = dN[:, i, :]
            sn[0] = acc(sn[0], ui[:, 0] + d[:, 0])
            sn[1] = acc(sn[1], ui[:, 1] + d[:, 1])
            sn[5] = acc(sn[5], 0.5 * ui[:, 0] * d[:, 1] + 0.5 * ui[:, 1] * d[:, 0])
            if dim > 2:
                sn[2] = acc(sn[2], ui[:, 2] + d[:, 2])
                sn[3] = acc(sn[3], 0.5 * ui[:, 1] * d[:, 2] + 0.5 * ui[:, 2] * d[:, 1])
                sn[4] = acc(sn[4], 0.5 * ui[:, 0] * d[:, 2] + 0.5 * ui[:, 2] * d[:, 0])
        if dim == 2:
            ps = (-nu * (sn[0] + sn[1]).astype(np.float64) / (1. - nu)).astype(F32)
            sn[2] = np.where(pstrain, ps, sn[2])
        tr = (sn[0] + sn[1]) + sn[2]
        f = lambda v: v.astype(np.float64)
        ss = [None] * 6
        ss[0] = (lam * f(tr) + 2. * mu * f(sn[0])).astype(F32)
        ss[5] = (2. * mu * f(sn[5])).astype(F32)
        ss[4] = (2. * mu * f(sn[4])).astype(F32)
        ss[1] = (lam * f(tr) + 2. * mu * f(sn[1])).astype(F32)
        ss[3] = (2. * mu * f(sn[3])).astype(F32)
        ss[2] = (lam * f(tr) + 2. * mu * f(sn[2])).astype(F32)
        if dim == 2:
            ss[2] = np.where(pstrain, ss[2], (nu * f(ss[0] + ss[1])).astype(F32))
        strain[:, q] = np.stack(sn, 1)
        stress[:, q] = np.stack(ss, 1)
    return xq.reshape(-1, 3), strain.reshape(-1, 6), stress.reshape(-1, 6)
