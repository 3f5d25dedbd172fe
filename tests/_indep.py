"""Independent formulations used to pin the oracle (numpy/scipy only).

None of these retype the oracle's loops: each computes the same quantity a
different way (tensor rotation instead of a Bond matrix, 2x2x2 Gauss instead of
3x3x3 Simpson, global sparse assembly instead of a matrix-free stencil, the
total-field element energy instead of the b/C split).
"""
import numpy as np

# Brick local node positions: node 1 (0,0,0), 2 (1,0,0), 3 (1,1,0), 4 (0,1,0),
# 5 (0,0,1), 6 (1,0,1), 7 (1,1,1), 8 (0,1,1) (src/elas3dxtal_pcg.f90:639, 740-765).
POS = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])

# Physical Voigt order xx, yy, zz, yz, xz, xy (engineering shear).
VOIGT_STD = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]
# Swap of Voigt rows/cols 4 and 5 (0-based 3 and 4).
P45 = np.eye(6)[[0, 1, 2, 4, 3, 5]]


def cubic_tensor(c11, c12, c44):
    c4 = np.zeros((3, 3, 3, 3))
    for i in range(3):
        for j in range(3):
            c4[i, i, j, j] = c12
        c4[i, i, i, i] = c11
    for i in range(3):
        for j in range(3):
            if i != j:
                c4[i, j, i, j] = c44
                c4[i, j, j, i] = c44
    return c4


def voigt(c4, order=VOIGT_STD):
    out = np.zeros((6, 6))
    for a, (i, j) in enumerate(order):
        for b, (k, l) in enumerate(order):
            out[a, b] = c4[i, j, k, l]
    return out


def rotate_tensor(c4, q):
    """C_lab[ijkl] = Q[pi] Q[qj] Q[rk] Q[sl] C[pqrs]  (Q maps lab to crystal coordinates)."""
    t = np.tensordot(q, c4, axes=([0], [0]))           # i q r s
    t = np.tensordot(q, t, axes=([0], [1]))            # j i r s
    t = np.tensordot(q, t, axes=([0], [2]))            # k j i s
    t = np.tensordot(q, t, axes=([0], [3]))            # l k j i
    return np.transpose(t, (3, 2, 1, 0))


def shape_grads(x, y, z):
    """dN_a/d(x,y,z) for the 8 trilinear shape functions at (x,y,z) in [0,1]^3."""
    pt = np.array([x, y, z])
    g = np.zeros((8, 3))
    for a in range(8):
        f = np.where(POS[a] == 1, pt, 1.0 - pt)
        df = np.where(POS[a] == 1, 1.0, -1.0)
        g[a, 0] = df[0] * f[1] * f[2]
        g[a, 1] = f[0] * df[1] * f[2]
        g[a, 2] = f[0] * f[1] * df[2]
    return g


def b_matrix_std(x, y, z):
    """6 x 24 strain-displacement matrix, physical order xx,yy,zz,yz,xz,xy; column a*3+c."""
    g = shape_grads(x, y, z)
    bm = np.zeros((6, 24))
    for a in range(8):
        gx, gy, gz = g[a]
        bm[0, 3 * a + 0] = gx
        bm[1, 3 * a + 1] = gy
        bm[2, 3 * a + 2] = gz
        bm[3, 3 * a + 1] = gz; bm[3, 3 * a + 2] = gy
        bm[4, 3 * a + 0] = gz; bm[4, 3 * a + 2] = gx
        bm[5, 3 * a + 0] = gy; bm[5, 3 * a + 1] = gx
    return bm


def element_stiffness_gauss(d_std):
    """24 x 24 element matrix by 2x2x2 Gauss quadrature (exact for the trilinear brick)."""
    gp = [0.5 - 0.5 / np.sqrt(3.0), 0.5 + 0.5 / np.sqrt(3.0)]
    k = np.zeros((24, 24))
    for x in gp:
        for y in gp:
            for z in gp:
                bm = b_matrix_std(x, y, z)
                k += 0.125 * bm.T @ d_std @ bm
    return k


def node_index(shape, i, j, k):
    nx, ny, nz = shape
    return (i % nx) + nx * ((j % ny) + ny * (k % nz))


def assemble_dense(shape, pix, kmats):
    """Global 3ns x 3ns matrix; dof (component c, node m) -> c*ns + m; kmats[p] is 24x24."""
    nx, ny, nz = shape
    ns = nx * ny * nz
    kg = np.zeros((3 * ns, 3 * ns))
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                e = node_index(shape, i, j, k)
                ke = kmats[pix[e] - 1]
                dofs = []
                for a in range(8):
                    nd = node_index(shape, i + POS[a, 0], j + POS[a, 1], k + POS[a, 2])
                    dofs += [nd, ns + nd, 2 * ns + nd]
                dofs = np.array(dofs)
                kg[np.ix_(dofs, dofs)] += ke
    return kg


def strain_tensor(e6):
    exx, eyy, ezz, exz, eyz, exy = e6
    return np.array([[exx, exy, exz], [exy, eyy, eyz], [exz, eyz, ezz]])


def voxel_jumps(shape, e6):
    """delta_e(a) = E (x_unwrapped - x_wrapped): the macro-strain displacement jump seen
    by local node a of the voxel at corner (i,j,k) when its node wraps periodically."""
    nx, ny, nz = shape
    et = strain_tensor(e6)
    dims = np.array([nx, ny, nz])
    out = {}
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                corner = np.array([i, j, k])
                shift = (corner + POS >= dims) * dims          # (8, 3)
                out[(i, j, k)] = shift @ et.T                  # (8, 3): et @ shift_a
    return out


def element_energy_total(shape, pix, kmats, u, e6):
    """U = 1/2 sum_e (u_e + delta_e)^T K_e (u_e + delta_e), u as (3, ns) planes."""
    nx, ny, nz = shape
    ns = nx * ny * nz
    jumps = voxel_jumps(shape, e6)
    uu = u.reshape(3, ns)
    tot = 0.0
    for (i, j, k), dl in jumps.items():
        e = node_index(shape, i, j, k)
        ve = np.zeros(24)
        for a in range(8):
            nd = node_index(shape, i + POS[a, 0], j + POS[a, 1], k + POS[a, 2])
            ve[3 * a:3 * a + 3] = uu[:, nd] + dl[a]
        tot += 0.5 * ve @ kmats[pix[e] - 1] @ ve
    return tot


def rhs_from_jumps(shape, pix, kmats, e6):
    """b = sum_e scatter(K_e delta_e), C = 1/2 sum_e delta_e^T K_e delta_e."""
    nx, ny, nz = shape
    ns = nx * ny * nz
    b = np.zeros(3 * ns)
    cval = 0.0
    for (i, j, k), dl in voxel_jumps(shape, e6).items():
        e = node_index(shape, i, j, k)
        d = dl.reshape(24)
        if not np.any(d):
            continue
        ke = kmats[pix[e] - 1]
        f = ke @ d
        cval += 0.5 * d @ f
        for a in range(8):
            nd = node_index(shape, i + POS[a, 0], j + POS[a, 1], k + POS[a, 2])
            for c in range(3):
                b[c * ns + nd] += f[3 * a + c]
    return b, cval


def jump_constant_exact(shape, pix, kmats, e6):
    """C = 1/2 sum_e delta_e^T K_e delta_e summed exactly (math.fsum of the voxel terms),
    vectorised for large grids: the voxels are grouped by which of their coordinates sit
    on the last plane (the jump is E times the wrap shift of each local node), and each
    group's constant is per phase."""
    import math
    nx, ny, nz = shape
    dims = np.array([nx, ny, nz])
    et = strain_tensor(e6)
    p3 = np.asarray(pix).reshape(nz, ny, nx) - 1
    terms = []
    for fx in (0, 1):
        for fy in (0, 1):
            for fz in (0, 1):
                if not (fx or fy or fz):
                    continue
                corner = np.array([nx - 1 if fx else 0, ny - 1 if fy else 0, nz - 1 if fz else 0])
                shift = (corner + POS >= dims) * dims * np.array([fx, fy, fz])
                d = (shift @ et.T).reshape(24)
                cph = 0.5 * np.einsum("pij,i,j->p", kmats, d, d)
                sx = slice(nx - 1, nx) if fx else slice(0, nx - 1)
                sy = slice(ny - 1, ny) if fy else slice(0, ny - 1)
                sz = slice(nz - 1, nz) if fz else slice(0, nz - 1)
                terms.append(cph[p3[sz, sy, sx].ravel()])
    return math.fsum(np.concatenate(terms).tolist())


def dk_to_kmats(dk):
    """dk[p,a,i,b,j] -> per-phase 24x24 with index a*3+i."""
    return dk.reshape(dk.shape[0], 24, 24)


def random_rodrigues(rng, n):
    """Uniform SO(3) via Shoemake quaternions, Rodrigues = q_xyz / q_w."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    qw = np.sqrt(1 - u1) * np.sin(2 * np.pi * u2)
    qx = np.sqrt(1 - u1) * np.cos(2 * np.pi * u2)
    qy = np.sqrt(u1) * np.sin(2 * np.pi * u3)
    qz = np.sqrt(u1) * np.cos(2 * np.pi * u3)
    return np.stack([qx / qw, qy / qw, qz / qw], axis=1)


def block_jacobi_dense(kg, ns):
    """Inverses of the 3x3 diagonal blocks of a dense K (dof c*ns + m), numpy; a block
    with |det| < 1e-14 gets a zero inverse.  Returns (ns, 3, 3)."""
    out = np.zeros((ns, 3, 3))
    for m in range(ns):
        idx = [m, ns + m, 2 * ns + m]
        blk = kg[np.ix_(idx, idx)]
        if abs(np.linalg.det(blk)) >= 1e-14:
            out[m] = np.linalg.inv(blk)
    return out


def pcg_reference_dense(kg, minv, u, g, gg0, tol, ldemb, reset=True, test_first=False):
    """The PCG of dembx_OpenMP written from its text (src/elas3dxtal_pcg.f90:1527-2007)
    over a dense matrix: z = M^-1 g, h = z; per step hAh (break if |hAh| < 1e-14),
    lambda = gg/hAh, u -= lambda h, g -= lambda Ah, z = M^-1 g, rel_res = sqrt(z.g/gg0);
    if rel_res > 10 rel_res_prev (step > 1) h = z and count a reset, else
    h = z + (gg/gglast) h; then stop when rel_res < tol.
    The other programs' loops: minv None is the plain CG of elas3dxtal_cg.f90 (z = g,
    cg:949-1056); reset=False drops the reset and test_first=True stops before the
    direction update (sweep A_2_0:1514-1526 and plain CG cg:1049-1056).
    Returns dict(u, g, h, history, lstep, resets, status) with status 1 converged,
    2 curvature break, 3 iteration limit."""
    ns = kg.shape[0] // 3
    u, g = u.copy(), g.copy()

    def prec(v):
        if minv is None:
            return v.copy()
        w = v.reshape(3, ns).T                     # (ns, 3)
        return np.einsum("mij,mj->mi", minv, w).T.ravel()

    z = prec(g)
    h = z.copy()
    gg = float(z @ g)
    rel_prev = np.sqrt(gg / gg0)
    resets, hist, status, step = 0, [], 3, 0
    for step in range(1, ldemb + 1):
        ah = kg @ h
        hah = float(h @ ah)
        if abs(hah) < 1e-14:
            status = 2
            break
        lam = gg / hah
        u -= lam * h
        g -= lam * ah
        z = prec(g)
        gglast, gg = gg, float(z @ g)
        rel = np.sqrt(gg / gg0)
        hist.append(rel)
        if test_first and rel < tol:
            status = 1
            break
        if reset and rel > 10.0 * rel_prev and step > 1:
            h = z.copy()
            resets += 1
        else:
            h = z + (gg / gglast) * h
        rel_prev = rel
        if rel < tol:
            status = 1
            break
    return dict(u=u, g=g, h=h, history=np.array(hist), lstep=step, resets=resets, status=status)
