"""Test helpers: the tissue diffusion operator assembled on the host from the oracle's Q1
element matrices (oracle/fem.c, restating Cfem::set_amat_bmat, fem.cpp:628-661) and solved
with the oracle's Jacobi-PCG (cg_solve_csr).  Test infrastructure only."""
import ctypes as C

import numpy as np

import oracle_lib as O

# element node order (0,0), (hx,0), (hx,hy), (0,hy)
OX = (0, 1, 1, 0)
OY = (0, 0, 1, 1)


def element(hx, hy, sxx, sxy, syy):
    Me = np.zeros(16)
    Ke = np.zeros(16)
    O.lib().fem_q1_element(hx, hy, sxx, sxy, syy, O.ptr(Me), O.ptr(Ke))
    return Me.reshape(4, 4), Ke.reshape(4, 4)


def assemble(nx, ny, hx, hy, sxx, sxy, syy, dt):
    """Dense A = M/dt + K and B = M/dt of the (nx+1)x(ny+1)-node grid (node iy*(nx+1)+ix)."""
    nn = (nx + 1) * (ny + 1)
    A = np.zeros((nn, nn))
    Bm = np.zeros((nn, nn))
    for ey in range(ny):
        for ex in range(nx):
            e = ey * nx + ex
            Me, Ke = element(hx, hy, sxx[e], sxy[e], syy[e])
            idx = [(ey + OY[j]) * (nx + 1) + ex + OX[j] for j in range(4)]
            for j in range(4):
                for k in range(4):
                    A[idx[j], idx[k]] += Me[j, k] / dt + Ke[j, k]
                    Bm[idx[j], idx[k]] += Me[j, k] / dt
    return A, Bm


def to_csr(A):
    rows, cols = np.nonzero(A)
    order = np.lexsort((cols, rows))
    rows, cols = rows[order], cols[order]
    rowptr = np.zeros(A.shape[0] + 1, dtype=np.int32)
    np.add.at(rowptr, rows + 1, 1)
    rowptr = np.cumsum(rowptr).astype(np.int32)
    return rowptr, cols.astype(np.int32), np.ascontiguousarray(A[rows, cols])


def cg(A, b, rtol, maxit=1000):
    rowptr, col, val = to_csr(A)
    x = np.zeros(len(b))
    bb = np.ascontiguousarray(b, dtype=np.float64)
    it = O.lib().cg_solve_csr(len(b), rowptr.ctypes.data_as(C.POINTER(C.c_int)),
                              col.ctypes.data_as(C.POINTER(C.c_int)), O.ptr(val), O.ptr(bb), O.ptr(x),
                              rtol, maxit, 1)
    return x, it


def random_tensors(rng, ne, dl=0.12, dt_=0.03):
    """Fibre-rotated tensors D = R diag(dl, dt) R^T with random angles (positive definite)."""
    th = rng.uniform(0, np.pi, ne)
    c, s = np.cos(th), np.sin(th)
    sxx = dl * c * c + dt_ * s * s
    syy = dl * s * s + dt_ * c * c
    sxy = (dl - dt_) * c * s
    return sxx, sxy, syy


def assemble_csr(nx, ny, hx, hy, sxx, sxy, syy, dt):
    """A = M/dt + K and B = M/dt as CSR (rowptr, col, val) triples, for grids too large for
    the dense assembly.  The element stiffness is linear in the tensor, so K_e is formed from
    the oracle's element matrices of the three unit tensors."""
    import scipy.sparse as sp
    Me, Kxx = element(hx, hy, 1.0, 0.0, 0.0)
    _, Kxy = element(hx, hy, 0.0, 1.0, 0.0)
    _, Kyy = element(hx, hy, 0.0, 0.0, 1.0)
    ne = nx * ny
    ex = np.tile(np.arange(nx), ny)
    ey = np.repeat(np.arange(ny), nx)
    idx = np.stack([(ey + OY[j]) * (nx + 1) + ex + OX[j] for j in range(4)], axis=1)   # ne x 4
    rows = np.repeat(idx, 4, axis=1).ravel()
    cols = np.tile(idx, (1, 4)).ravel()
    sxx, sxy, syy = (np.broadcast_to(np.asarray(v, np.float64), (ne,)) for v in (sxx, sxy, syy))
    Ke = (sxx[:, None] * Kxx.ravel()[None, :] + sxy[:, None] * Kxy.ravel()[None, :]
          + syy[:, None] * Kyy.ravel()[None, :])
    Mv = np.broadcast_to(Me.ravel()[None, :] / dt, (ne, 16))
    nn = (nx + 1) * (ny + 1)
    out = []
    for vals in (Mv + Ke, Mv):
        m = sp.coo_matrix((np.ascontiguousarray(vals).ravel(), (rows, cols)), shape=(nn, nn)).tocsr()
        m.sort_indices()
        out.append((m.indptr.astype(np.int32), m.indices.astype(np.int32), np.ascontiguousarray(m.data)))
    return out[0], out[1]


def cg_csr(csr, b, rtol, maxit=1000):
    rowptr, col, val = csr
    x = np.zeros(len(b))
    bb = np.ascontiguousarray(b, dtype=np.float64)
    it = O.lib().cg_solve_csr(len(b), rowptr.ctypes.data_as(C.POINTER(C.c_int)),
                              col.ctypes.data_as(C.POINTER(C.c_int)), O.ptr(val), O.ptr(bb), O.ptr(x),
                              rtol, maxit, 1)
    return x, it


def q4_element(xe, ye, sxx, sxy, syy):
    """The oracle's element matrices on any quadrilateral (nodes counter-clockwise)."""
    Me = np.zeros(16)
    Ke = np.zeros(16)
    x = np.ascontiguousarray(xe, dtype=np.float64)
    y = np.ascontiguousarray(ye, dtype=np.float64)
    O.lib().fem_q4_element(O.ptr(x), O.ptr(y), sxx, sxy, syy, O.ptr(Me), O.ptr(Ke))
    return Me.reshape(4, 4), Ke.reshape(4, 4)


def assemble_mesh(xy, conn, sxx, sxy, syy, dt):
    """Dense A = M/dt + K and B = M/dt of an S4 mesh (xy: nn x 2, conn: ne x 4)."""
    nn = len(xy)
    A = np.zeros((nn, nn))
    Bm = np.zeros((nn, nn))
    for e, nd in enumerate(conn):
        Me, Ke = q4_element(xy[nd, 0], xy[nd, 1], sxx[e], sxy[e], syy[e])
        for j in range(4):
            for k in range(4):
                A[nd[j], nd[k]] += Me[j, k] / dt + Ke[j, k]
                Bm[nd[j], nd[k]] += Me[j, k] / dt
    return A, Bm


def grid_mesh(nx, ny, hx, hy):
    """The structured grid as an S4 mesh with the grid's node and element numbering."""
    xs, ys = np.meshgrid(np.arange(nx + 1) * hx, np.arange(ny + 1) * hy)
    xyz = np.stack([xs.ravel(), ys.ravel(), np.zeros(xs.size)], axis=1)
    conn = [[(ey + OY[j]) * (nx + 1) + ex + OX[j] for j in range(4)] for ey in range(ny) for ex in range(nx)]
    return xyz, np.array(conn, dtype=np.int32)


def distorted_mesh(nx, ny, h, rng, amp=0.25):
    """A grid whose interior nodes are moved by up to amp*h: every element stays convex."""
    xyz, conn = grid_mesh(nx, ny, h, h)
    ix = np.tile(np.arange(nx + 1), ny + 1)
    iy = np.repeat(np.arange(ny + 1), nx + 1)
    inner = (ix > 0) & (ix < nx) & (iy > 0) & (iy < ny)
    xyz[inner, 0] += rng.uniform(-amp, amp, inner.sum()) * h
    xyz[inner, 1] += rng.uniform(-amp, amp, inner.sum()) * h
    return xyz, conn


def write_inp(path, xyz, conn, targets=(1, 1, 1), celltype=None, etype="S4"):
    """A mesh file in the section layout Cmesh::loadInp reads (mesh.cpp:59-183)."""
    with open(path, "w") as f:
        f.write("** synthetic mesh\n")
        f.write("CML\n")
        f.write(f"{len(xyz)} {len(conn)} {targets[0]} {targets[1]} {targets[2]}\n")
        f.write("*NODE\n")
        for i, p in enumerate(xyz):
            f.write(f"{i + 1}, {p[0]:.17g}, {p[1]:.17g}, {p[2]:.17g}\n")
        f.write("**HWCOLOR COMP 1 4\n")
        f.write(f"*ELEMENT,TYPE={etype},ELSET=tissue\n")
        for e, nd in enumerate(conn):
            f.write(f"{e + 1}, " + ", ".join(str(int(v) + 1) for v in nd) + "\n")
        if celltype is not None:
            f.write("CELLTYPE\n")
            for c in celltype:
                f.write(f"{int(c)}\n")
        f.write("eof\n")


# ---------------------------------------------------------------- 3-D meshes

def hex8_element(X, D):
    """The oracle's trilinear hexahedron: X 8 x 3 (the reference's node order), D = (xx, xy,
    xz, yy, yz, zz)."""
    Me = np.zeros(64)
    Ke = np.zeros(64)
    x = np.ascontiguousarray(X, dtype=np.float64)
    d = np.ascontiguousarray(D, dtype=np.float64)
    O.lib().fem_hex8_element(O.ptr(x), O.ptr(d), O.ptr(Me), O.ptr(Ke))
    return Me.reshape(8, 8), Ke.reshape(8, 8)


def tet4_element(X, D):
    Me = np.zeros(16)
    Ke = np.zeros(16)
    x = np.ascontiguousarray(X, dtype=np.float64)
    d = np.ascontiguousarray(D, dtype=np.float64)
    O.lib().fem_tet4_element(O.ptr(x), O.ptr(d), O.ptr(Me), O.ptr(Ke))
    return Me.reshape(4, 4), Ke.reshape(4, 4)


def assemble_mesh3(xyz, conn, D, dt):
    """Dense A = M/dt + K and B = M/dt of a hexahedral (8 nodes) or tetrahedral (4) mesh."""
    nn = len(xyz)
    A = np.zeros((nn, nn))
    Bm = np.zeros((nn, nn))
    for e, nd in enumerate(conn):
        Me, Ke = (hex8_element if len(nd) == 8 else tet4_element)(xyz[nd], D[e])
        n = len(nd)
        for j in range(n):
            for k in range(n):
                A[nd[j], nd[k]] += Me[j, k] / dt + Ke[j, k]
                Bm[nd[j], nd[k]] += Me[j, k] / dt
    return A, Bm


def box_hex_mesh(nx, ny, nz, hx, hy, hz):
    """A box of nx x ny x nz hexahedra; node (ix, iy, iz) at index (iz*(ny+1) + iy)*(nx+1) + ix,
    elements in the reference's node order."""
    xs, ys, zs = np.meshgrid(np.arange(nx + 1) * hx, np.arange(ny + 1) * hy, np.arange(nz + 1) * hz,
                             indexing="ij")
    xyz = np.stack([xs.transpose(2, 1, 0).ravel(), ys.transpose(2, 1, 0).ravel(), zs.transpose(2, 1, 0).ravel()],
                   axis=1)

    def node(ix, iy, iz):
        return (iz * (ny + 1) + iy) * (nx + 1) + ix
    conn = []
    for ez in range(nz):
        for ey in range(ny):
            for ex in range(nx):
                bot = [node(ex + OX[j], ey + OY[j], ez) for j in range(4)]
                top = [node(ex + OX[j], ey + OY[j], ez + 1) for j in range(4)]
                conn.append(bot + top)
    return xyz, np.array(conn, dtype=np.int32)


def distorted_box(nx, ny, nz, h, rng, amp=0.2):
    """A box of hexahedra whose interior nodes are moved by up to amp*h in each direction."""
    xyz, conn = box_hex_mesh(nx, ny, nz, h, h, h)
    idx = np.arange(len(xyz))
    ix, iy, iz = idx % (nx + 1), (idx // (nx + 1)) % (ny + 1), idx // ((nx + 1) * (ny + 1))
    inner = (ix > 0) & (ix < nx) & (iy > 0) & (iy < ny) & (iz > 0) & (iz < nz)
    xyz[inner] += rng.uniform(-amp, amp, (inner.sum(), 3)) * h
    return xyz, conn


# the six tetrahedra of a cube along its (0,0,0)-(1,1,1) diagonal (local hexahedron node
# numbers): neighbouring cubes share their face diagonals, so the mesh is conforming
_KUHN = [(0, 1, 2, 6), (0, 2, 3, 6), (0, 3, 7, 6), (0, 7, 4, 6), (0, 4, 5, 6), (0, 5, 1, 6)]


def tet_mesh_from_hex(xyz, conn):
    """Split every hexahedron into six tetrahedra, each with a positive orientation."""
    tets = []
    for nd in conn:
        for q in _KUHN:
            t = [nd[k] for k in q]
            p = xyz[t]
            if np.linalg.det(np.stack([p[1] - p[0], p[2] - p[0], p[3] - p[0]])) < 0:
                t[2], t[3] = t[3], t[2]
            tets.append(t)
    return np.array(tets, dtype=np.int32)


def random_tensors3(rng, ne, dl=0.12, dt_=0.03):
    """Fibre tensors D = dt I + (dl - dt) f f^T with random unit fibre directions f."""
    f = rng.normal(size=(ne, 3))
    f /= np.linalg.norm(f, axis=1, keepdims=True)
    D = dt_ * np.eye(3)[None] + (dl - dt_) * f[:, :, None] * f[:, None, :]
    return np.stack([D[:, 0, 0], D[:, 0, 1], D[:, 0, 2], D[:, 1, 1], D[:, 1, 2], D[:, 2, 2]], axis=1)
