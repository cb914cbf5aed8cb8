"""Explicit interface lists for mmf_mesh_create (the general MeshGeometricalInfo,
mesh_info.hpp:31-105): owner / neighbour / area / owner-outward normal per
face and volume / size / centroid per cell, as bitpit's VolOctree hands them
to the reference's interface loop (euler.cpp:153-247).

  cartesian(dim, n, origin, h, side_bc, side_info)
      a uniform grid as a face list: faces axis by axis (x, y, z), within an
      axis plane-major (k, then j, then i); an interior face is owned by its
      lower cell with normal +e_a, a border face by its cell with the outward
      normal; a periodic axis has its wrap face first in each row (owner the
      last cell, neighbour the first).
  refined(dim, n, origin, h, box, side_bc, side_info)
      the same grid with every cell whose centroid lies in 'box' split into
      2^dim children (a 2:1 octree level): a coarse cell next to refined
      cells has 2^(dim-1) hanging faces on that side, each of the fine area.
  morton_order(mesh)
      the faces reordered along a Morton (Z-order) curve of their centroids.

Plain numpy; nothing here runs on the GPU.
"""
import numpy as np

NF = 5


class FaceMesh:
    """Arrays of one explicit mesh (cells numbered 0..ncells-1)."""

    def __init__(self, dim, volume, size, centroid, owner, neigh, area, normal, face_centroid, face_bc, face_info):
        self.dim = int(dim)
        self.volume = np.ascontiguousarray(volume, dtype=np.float64)
        self.size = np.ascontiguousarray(size, dtype=np.float64)
        self.centroid = np.ascontiguousarray(centroid, dtype=np.float64).reshape(-1, 3)
        self.owner = np.ascontiguousarray(owner, dtype=np.int64)
        self.neigh = np.ascontiguousarray(neigh, dtype=np.int64)
        self.area = np.ascontiguousarray(area, dtype=np.float64)
        self.normal = np.ascontiguousarray(normal, dtype=np.float64).reshape(-1, 3)
        self.face_centroid = np.ascontiguousarray(face_centroid, dtype=np.float64).reshape(-1, 3)
        self.face_bc = np.ascontiguousarray(face_bc, dtype=np.int32)
        self.face_info = np.ascontiguousarray(face_info, dtype=np.float64).reshape(-1, NF)

    @property
    def ncells(self):
        return len(self.volume)

    @property
    def nfaces(self):
        return len(self.owner)

    @property
    def min_cell_size(self):
        return float(self.size.min())

    def permuted(self, order):
        """The same mesh with faces in the given order."""
        o = np.asarray(order)
        return FaceMesh(self.dim, self.volume, self.size, self.centroid, self.owner[o], self.neigh[o], self.area[o],
                        self.normal[o], self.face_centroid[o], self.face_bc[o], self.face_info[o])


def _side_arrays(side_bc, side_info):
    bc = [int(b) for b in side_bc] + [0] * (6 - len(side_bc))
    info = np.zeros((6, NF)) if side_info is None else np.asarray(side_info, dtype=np.float64).reshape(6, NF)
    return bc, info


def _leaf_faces(dim, nf, origin, hf, leaf, periodic, side_bc, side_info):
    """Faces between the leaves of a fine-level grid nf (leaf[fine cell] = the
    cell it belongs to), merged where both sides are the same pair of leaves
    along the same axis position.  Returns face arrays in the structured
    order of their first fine face."""
    n = [int(x) for x in nf] + [1] * (3 - len(nf))
    own, nei, ar, nrm, fc, fbc, finfo, key = [], [], [], [], [], [], [], []
    hA = hf * hf if dim == 3 else hf  # the closed forms' h*h / h (no pow)
    for a in range(dim):
        lim = list(n)
        lim[a] = n[a] if periodic[a] else n[a] + 1
        K, J, I = np.meshgrid(np.arange(lim[2]), np.arange(lim[1]), np.arange(lim[0]), indexing="ij")
        ijk = [I.ravel(), J.ravel(), K.ravel()]
        p = ijk[a]
        lo = [x.copy() for x in ijk]
        hi = [x.copy() for x in ijk]
        lo[a] = p - 1
        if periodic[a]:
            lo[a] = np.where(p == 0, n[a] - 1, p - 1)
        lo_ok = lo[a] >= 0
        hi_ok = hi[a] < n[a]

        def cell(q, ok):
            qq = [np.where(ok, x, 0) for x in q]
            return qq[0] + n[0] * (qq[1] + n[1] * qq[2])

        lo_leaf = np.where(lo_ok, leaf[cell(lo, lo_ok)], -1)
        hi_leaf = np.where(hi_ok, leaf[cell(hi, hi_ok)], -1)
        keep = lo_leaf != hi_leaf
        order_key = a * (1 << 60) + (ijk[0] + lim[0] * (ijk[1] + lim[1] * ijk[2]))
        # face centre of each fine face
        cen = np.zeros((len(p), 3))
        for d in range(3):
            cen[:, d] = origin[d] + (ijk[d] + 0.5) * hf if d < dim else origin[d]
        cen[:, a] = origin[a] + p * hf
        lo_l, hi_l, ok_k = lo_leaf[keep], hi_leaf[keep], order_key[keep]
        c_sel, p_sel = cen[keep], p[keep]
        # merge fine faces sharing (axis, position, lower leaf, upper leaf): one face
        tag = np.stack([p_sel, lo_l, hi_l], axis=1)
        _, first, inv, cnt = np.unique(tag, axis=0, return_index=True, return_inverse=True, return_counts=True)
        inv = inv.ravel()
        csum = np.zeros((len(first), 3))
        np.add.at(csum, inv, c_sel)
        lo_u, hi_u = lo_l[first], hi_l[first]
        border_lo = lo_u < 0
        border_hi = hi_u < 0
        owner = np.where(border_lo, hi_u, lo_u)
        neigh = np.where(border_lo | border_hi, -1, hi_u)
        nv = np.zeros((len(first), 3))
        nv[:, a] = np.where(border_lo, -1.0, 1.0)
        side = np.where(border_lo, 2 * a, np.where(border_hi, 2 * a + 1, -1))
        own.append(owner)
        nei.append(neigh)
        ar.append(cnt * hA)
        nrm.append(nv)
        fc.append(csum / cnt[:, None])
        fbc.append(np.where(side >= 0, np.array(side_bc)[np.maximum(side, 0)], -1))
        finfo.append(np.where((side >= 0)[:, None], side_info[np.maximum(side, 0)], 0.0))
        key.append(ok_k[first])
    key = np.concatenate(key)
    o = np.argsort(key, kind="stable")
    cat = lambda xs: np.concatenate(xs)[o]
    return cat(own), cat(nei), cat(ar), cat(nrm), cat(fc), cat(fbc), cat(finfo)


def cartesian(dim, n, origin, h, side_bc, side_info=None):
    """A uniform grid as an explicit face list (the oracle's and the structured
    kernels' face order)."""
    bc, info = _side_arrays(side_bc, side_info)
    n = [int(x) for x in n] + [1] * (3 - len(n))
    N = n[0] * n[1] * n[2]
    periodic = [a < dim and bc[2 * a] == 4 for a in range(3)]
    leaf = np.arange(N)
    own, nei, ar, nrm, fc, fbc, finfo = _leaf_faces(dim, n[:dim], origin, h, leaf, periodic, bc, info)
    K, J, I = np.meshgrid(np.arange(n[2]), np.arange(n[1]), np.arange(n[0]), indexing="ij")
    cen = np.zeros((N, 3))
    for d, q in enumerate((I.ravel(), J.ravel(), K.ravel())):
        cen[:, d] = origin[d] + (q + 0.5) * h if d < dim else origin[d]
    vol = np.full(N, h * h * h if dim == 3 else h * h)
    ar = np.full(len(own), h * h if dim == 3 else h)
    return FaceMesh(dim, vol, np.full(N, h), cen, own, nei, ar, nrm, fc, fbc, finfo)


def refined(dim, n, origin, h, box, side_bc, side_info=None):
    """A uniform grid of n^dim cells of size h with the cells whose centroid
    lies inside box = (lo[3], hi[3]) split into 2^dim children of size h/2."""
    bc, info = _side_arrays(side_bc, side_info)
    n = [int(x) for x in n] + [1] * (3 - len(n))
    periodic = [a < dim and bc[2 * a] == 4 for a in range(3)]
    nf = [2 * n[a] if a < dim else 1 for a in range(3)]
    hf = 0.5 * h
    Kf, Jf, If = np.meshgrid(np.arange(nf[2]), np.arange(nf[1]), np.arange(nf[0]), indexing="ij")
    fi = [If.ravel(), Jf.ravel(), Kf.ravel()]
    ci = [fi[d] // 2 if d < dim else fi[d] for d in range(3)]
    ccen = np.zeros((len(fi[0]), 3))
    for d in range(3):
        ccen[:, d] = origin[d] + (ci[d] + 0.5) * h if d < dim else origin[d]
    lo, hi = np.asarray(box[0], dtype=float), np.asarray(box[1], dtype=float)
    fine = np.all((ccen[:, :dim] >= lo[:dim]) & (ccen[:, :dim] <= hi[:dim]), axis=1)
    coarse_id = ci[0] + n[0] * (ci[1] + n[1] * ci[2])
    fine_id = np.arange(len(fi[0]))
    # a leaf key per fine-level cell; numbered in order of first appearance
    tag = np.where(fine, fine_id, -1 - coarse_id)
    _, first, leaf = np.unique(tag, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(len(order))
    leaf = rank[leaf.ravel()]
    nleaf = len(first)
    is_fine = np.zeros(nleaf, dtype=bool)
    is_fine[leaf[fine]] = True
    size = np.where(is_fine, hf, h)
    vol = size * size * size if dim == 3 else size * size
    cen = np.zeros((nleaf, 3))
    fcen = np.zeros((len(fi[0]), 3))
    for d in range(3):
        fcen[:, d] = origin[d] + (fi[d] + 0.5) * hf if d < dim else origin[d]
    cnt = np.bincount(leaf, minlength=nleaf)
    for d in range(3):
        cen[:, d] = np.bincount(leaf, weights=fcen[:, d], minlength=nleaf) / cnt
    own, nei, ar, nrm, fc, fbc, finfo = _leaf_faces(dim, nf[:dim], origin, hf, leaf, periodic, bc, info)
    return FaceMesh(dim, vol, size, cen, own, nei, ar, nrm, fc, fbc, finfo)


def morton_order(mesh, bits=16):
    """Face permutation along a Morton curve of the face centroids."""
    c = mesh.face_centroid[:, :mesh.dim]
    lo, hi = c.min(0), c.max(0)
    q = np.floor((c - lo) / np.maximum(hi - lo, 1e-300) * ((1 << bits) - 1)).astype(np.int64)
    code = np.zeros(len(c), dtype=np.int64)
    for b in range(bits):
        for d in range(mesh.dim):
            code |= ((q[:, d] >> b) & 1) << (b * mesh.dim + d)
    return np.argsort(code, kind="stable")
