"""Writes mesh_tables.json: element order, face-neighbour lists, neighbour
faces, ajsurf modes and orientation codes of small meshes, derived from
nmesh's own rules without the oracle's geometric neighbour search.

Run from the repository root:  python tests/golden/make_mesh_tables.py

Rules restated (paths relative to the nmesh source tree):

- Element locations: an element is (patch, loc) with one octant digit per
  level; bit d of a digit is set for the upper half in direction d
  (amr_set_elm0_bbox, src/main/amr/connections.c:1160-1206).  Refinement
  replaces an element by its 8 children, digits 0..7, in place in the element
  list, and eids follow the list (update_elm_eid_dt,
  src/main/amr/storage.c:1305-1352): patches in order, depth-first octant
  order inside each.
- Patches: BoxMesh is one patch; the Shell is the 6 cubed-sphere patches in
  dom order 0..5 (add_N_CubedSphere_pats,
  src/main/coordinates/setup_CubedSpheres.c:653-686), dom = 2 dir + (sign > 0),
  with the map xyz_of_lamAB_CubSph (coordtrans_CubedSphere.c:38-120).
- Face neighbours inside a patch: the same-level location across the face
  (connections_get_nbloc_InsidePat, connections.c:313), then its ancestor
  that is a leaf, or else its leaf descendants on the facing side in
  depth-first child order 0..7 (amr_make_elms_on_eloc_face_list,
  connections.c:1549-1660; connections_ijk_is_at_parentface, :193).
- Across a patch face: the leaves of the other patch on its facing patch
  face, in the same depth-first order from that patch's root, keeping those
  that share face points away from edges (amr_make_patchface_fnb_list,
  connections.c:1662-1730; find_elmfacepoints_in_nbface, :388-451).
- The neighbour's face: f^1 inside a patch (amr_get_nbface, connections.c:
  1114-1157), the other patch face across.
- ajsurf mode (set_ajsurf_vl, src/main/amr/surface.c:799-1080): 0 no
  neighbour, 1 one neighbour of the same patch and level (alias), 2 one
  neighbour of another patch whose face has only this element (oriented
  copy), 3 otherwise (interpolation).  All elements here have one n.
- Orientation code intrch*4 + rev1*2 + rev2 of the copy
  (copy_ajsurf_from_nbsurf0, surface.c:1361-1466): my face point (I1, I2)
  along (Dir1_norm, Dir2_norm) takes the neighbour's point (o1, o2) with
  o = (I1, I2) or (I2, I1) under intrch, each reversed when rev1 / rev2.
  The code is the one under which every corner of my face lands on the
  neighbour corner at the same physical position.
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def at_parentface(ijk, face):
    b = (ijk >> (face // 2)) & 1
    return b == (face & 1)


class Forest:
    def __init__(self, kind, luni, dout=8.0, rin=2.0, rout=20.0):
        self.kind, self.dout, self.rin, self.rout = kind, dout, rin, rout
        npat = 1 if kind == "box" else 6
        self.elms = []
        for p in range(npat):
            locs = [""]
            for _ in range(luni):
                locs = [l + str(c) for l in locs for c in range(8)]
            self.elms += [(p, l) for l in locs]

    def refine(self, flags):
        out = []
        for (p, l), fl in zip(self.elms, flags):
            out += [(p, l + str(c)) for c in range(8)] if fl else [(p, l)]
        self.elms = out

    # geometry of locations
    def pat_bbox(self, p):
        if self.kind == "box":
            return np.array([-self.dout, self.dout] * 3)
        return np.array([0.0, 1.0, -1.0, 1.0, -1.0, 1.0])

    def bbox(self, p, loc):
        bb = self.pat_bbox(p).astype(float)
        for ch in loc:
            c = int(ch)
            for d in range(3):
                half = 0.5 * (bb[2 * d + 1] - bb[2 * d])
                if (c >> d) & 1:
                    bb[2 * d] = bb[2 * d + 1] - half
                else:
                    bb[2 * d + 1] = bb[2 * d] + half
        return bb

    def xyz(self, p, X):
        if self.kind == "box":
            return np.asarray(X, float)
        lam, A, B = X
        d, pm = p // 2, 2 * (p % 2) - 1
        q = np.sqrt(1.0 + A * A + B * B)
        a0, a1 = pm * self.rin / q, pm * self.rout / q
        r = (a1 - a0) * lam + a0
        return np.array([[r, A * r, B * r], [A * r, r, B * r], [B * r, A * r, r]][d])

    def X_of(self, p, x):
        """patch coordinates of x in patch p (None if x is not in it)"""
        if self.kind == "box":
            return np.asarray(x, float)
        d, pm = p // 2, 2 * (p % 2) - 1
        ia, ib = [(1, 2), (0, 2), (1, 0)][d]
        r = x[d]
        if r * pm <= 0:
            return None
        A, B = x[ia] / r, x[ib] / r
        q = np.sqrt(1.0 + A * A + B * B)
        a0, a1 = pm * self.rin / q, pm * self.rout / q
        X = np.array([(r - a0) / (a1 - a0), A, B])
        if (X < self.pat_bbox(p)[0::2] - 1e-12).any() or (X > self.pat_bbox(p)[1::2] + 1e-12).any():
            return None
        return X

    def face_point(self, p, bb, f, s1, s2):
        """physical point of face f of box bb at in-plane fractions (s1, s2)
        along (Dir1_norm, Dir2_norm)"""
        d = f // 2
        d1, d2 = (1 if d == 0 else 0), (1 if d == 2 else 2)
        X = np.zeros(3)
        X[d] = bb[f]
        X[d1] = bb[2 * d1] + s1 * (bb[2 * d1 + 1] - bb[2 * d1])
        X[d2] = bb[2 * d2] + s2 * (bb[2 * d2 + 1] - bb[2 * d2])
        return self.xyz(p, X)

    # neighbours
    def leaves_on_face(self, p, loc, s_f, leaves):
        if (p, loc) in leaves:
            return [(p, loc)]
        if len(loc) > 12:
            return []
        out = []
        for c in range(8):
            if at_parentface(c, s_f):
                out += self.leaves_on_face(p, loc + str(c), s_f, leaves)
        return out

    def patch_face_partner(self, p, f):
        """(other patch, its face) touching patch face f of p, or None"""
        if self.kind == "box" or f < 2:
            return None
        x = self.face_point(p, self.pat_bbox(p), f, 0.5, 0.5)
        for p2 in range(6):
            if p2 == p:
                continue
            X = self.X_of(p2, x)
            if X is None:
                continue
            bb2 = self.pat_bbox(p2)
            for f2 in range(6):
                if abs(X[f2 // 2] - bb2[f2]) < 1e-12:
                    return p2, f2
        return None

    def touches(self, p, bb, f, p2, bb2, f2):
        """faces share a piece of surface (not only an edge)"""
        pts = [self.X_of(p2, self.face_point(p, bb, f, s1, s2)) for s1 in (0, 1) for s2 in (0, 1)]
        d2 = f2 // 2
        ax = [a for a in range(3) if a != d2]
        lo = [min(P[a] for P in pts) for a in ax]
        hi = [max(P[a] for P in pts) for a in ax]
        ov = [min(hi[i], bb2[2 * a + 1]) - max(lo[i], bb2[2 * a]) for i, a in enumerate(ax)]
        return min(ov) > 1e-9

    def tables(self):
        leaves = set(self.elms)
        index = {e: i for i, e in enumerate(self.elms)}
        out = []
        for p, loc in self.elms:
            l = len(loc)
            bb = self.bbox(p, loc)
            I = [sum(((int(c) >> d) & 1) << (l - 1 - m) for m, c in enumerate(loc)) for d in range(3)]
            faces = []
            for f in range(6):
                d, s = f // 2, (1 if f & 1 else -1)
                J = list(I)
                J[d] += s
                if 0 <= J[d] < (1 << l):
                    nbloc = "".join(str(sum(((J[dd] >> (l - 1 - m)) & 1) << dd for dd in range(3)))
                                    for m in range(l))
                    found = None
                    for ll in range(0, l + 1):
                        if (p, nbloc[:ll]) in leaves:
                            found = [(p, nbloc[:ll])]
                            break
                    nbs = found or self.leaves_on_face(p, nbloc, f ^ 1, leaves)
                    faces.append([[index[e] for e in nbs], [f ^ 1] * len(nbs)])
                    continue
                part = self.patch_face_partner(p, f)
                if part is None:
                    faces.append([[], []])
                    continue
                p2, f2 = part
                nbs = [e for e in self.leaves_on_face(p2, "", f2, leaves)
                       if self.touches(p, bb, f, p2, self.bbox(*e), f2)]
                faces.append([[index[e] for e in nbs], [f2] * len(nbs)])
            out.append({"p": p, "loc": loc, "bbox": bb.tolist(), "faces": faces})
        # ajsurf modes and orientation codes
        for e, E in enumerate(out):
            modes, codes = [], []
            for f in range(6):
                nbs, nbf = E["faces"][f]
                mode, code = (0 if not nbs else 3), 0
                if len(nbs) == 1:
                    N = out[nbs[0]]
                    if N["p"] == E["p"]:
                        mode = 1 if len(N["loc"]) == len(E["loc"]) else 3
                    elif len(N["faces"][nbf[0]][0]) == 1:
                        code = self.orientation(E, f, N, nbf[0])
                        mode = 2 if code >= 0 else 3
                        code = max(code, 0)
                modes.append(mode)
                codes.append(code)
            E["ajmode"], E["orient"] = modes, codes
        return out

    def orientation(self, E, f, N, nf):
        bbE, bbN = np.array(E["bbox"]), np.array(N["bbox"])
        corners = [(a, b) for a in (0, 1) for b in (0, 1)]
        where = {}
        for a, b in corners:
            x = self.face_point(E["p"], bbE, f, a, b)
            hit = [(c, d) for c, d in corners
                   if np.abs(self.face_point(N["p"], bbN, nf, c, d) - x).max() < 1e-9 * max(1, np.abs(x).max())]
            if len(hit) != 1:
                return -1
            where[(a, b)] = hit[0]
        for code in range(8):
            intrch, rev1, rev2 = code >> 2, (code >> 1) & 1, code & 1
            ok = True
            for (a, b), (c, d) in where.items():
                o1, o2 = (b, a) if intrch else (a, b)
                o1 = 1 - o1 if rev1 else o1
                o2 = 1 - o2 if rev2 else o2
                ok = ok and (o1, o2) == (c, d)
            if ok:
                return code
        return -1


def cases():
    out = {}
    f = Forest("shell", 1)
    out["shell_luni1"] = {"kind": "shell", "luni": 1, "rin": 2.0, "rout": 20.0, "refine": [],
                          "elms": f.tables()}
    f = Forest("box", 1)
    flags = [1 if i == 0 else 0 for i in range(len(f.elms))]
    f.refine(flags)
    out["box_luni1_refine0"] = {"kind": "box", "luni": 1, "dout": 8.0, "refine": [flags], "elms": f.tables()}
    f = Forest("box", 1)
    f.refine(flags)
    flags2 = [1 if i in (3, 9) else 0 for i in range(len(f.elms))]
    f.refine(flags2)
    out["box_two_levels"] = {"kind": "box", "luni": 1, "dout": 8.0, "refine": [flags, flags2],
                             "elms": f.tables()}
    f = Forest("shell", 1)
    flags = [1 if i in (9, 26) else 0 for i in range(len(f.elms))]
    f.refine(flags)
    out["shell_luni1_refined"] = {"kind": "shell", "luni": 1, "rin": 2.0, "rout": 20.0, "refine": [flags],
                                  "elms": f.tables()}
    return out


if __name__ == "__main__":
    with open(os.path.join(HERE, "mesh_tables.json"), "w") as fh:
        json.dump(cases(), fh, separators=(",", ":"))
