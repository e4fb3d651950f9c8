"""Build CSGFoundry-layout geometry and SSim-layout physics tables with numpy.

The arrays produced here are byte-compatible with what eic-opticks persists and
uploads, so the same engine entry points accept either a reference geometry
folder (see ``oxr.foundry.load_foundry``) or a synthetic scene from
``oxr.scenes``:

* CSGNode 64 B, level-order complete binary trees        CSG/CSGNode.h:67-98,196
* CSGPrim 64 B (numNode, nodeOffset, ..., AABB)          CSG/CSGPrim.h:20-60
* CSGSolid 48 B (label, numPrim, primOffset, ...)        CSG/CSGSolid.h:37-53
* qat4 instance transforms, identity in the .w column    sysrap/sqat4.h:345-389
* bnd (nb,4,2,nw,4) and optical (nb*4,4)                 sysrap/sstandard.h:311-470
* surface payload (detect, absorb, specular, diffuse)    u4/U4SurfaceArray.h:159-258

Transforms follow the reference's row-vector convention: a point p (row) maps
to p @ M[:3,:3] + M[3,:3], so the translation lives in row 3.
"""
from __future__ import annotations

import numpy as np

# ---- OpticksCSG_t typecodes (sysrap/OpticksCSG.h) -------------------------
CSG_ZERO = 0
CSG_UNION = 1
CSG_INTERSECTION = 2
CSG_DIFFERENCE = 3
CSG_CONTIGUOUS = 11
CSG_DISCONTIGUOUS = 12
CSG_OVERLAP = 13
CSG_SPHERE = 101
CSG_ZSPHERE = 103
CSG_CYLINDER = 105
CSG_CONE = 108
CSG_BOX3 = 110
CSG_CONVEXPOLYHEDRON = 112
CSG_HYPERBOLOID = 117
CSG_PHICUT = 121
CSG_HALFSPACE = 125

# ---- smatsur "ems" enum (sysrap/smatsur.h:6-14) ----------------------------
SMATSUR_MATERIAL = 0
SMATSUR_NOSURFACE = 1
SMATSUR_SURFACE = 2
SMATSUR_SENSOR_A = 3

# sdomain: 60..820 nm in 1 nm steps (sysrap/sdomain.h)
WL_MIN, WL_MAX, WL_STEP = 60.0, 820.0, 1.0
NUM_WL = int(round((WL_MAX - WL_MIN) / WL_STEP)) + 1     # 761
C_LIGHT = 299.792458                                      # mm/ns
HC_EVNM = 1239.84198                                      # h*c in eV*nm


def domain():
    return WL_MIN + WL_STEP * np.arange(NUM_WL, dtype=np.float64)


# ---------------------------------------------------------------------------
# transforms
# ---------------------------------------------------------------------------
def translate(x, y, z):
    m = np.eye(4)
    m[3, :3] = (x, y, z)
    return m


def scale(sx, sy, sz):
    return np.diag([sx, sy, sz, 1.0])


def rotate_z_to(d):
    """Rotation (row convention) that maps local +z onto unit vector d."""
    d = np.asarray(d, dtype=np.float64)
    d = d / np.linalg.norm(d)
    z = np.array([0.0, 0.0, 1.0])
    v = np.cross(z, d)
    c = float(np.dot(z, d))
    if np.linalg.norm(v) < 1e-12:
        r = np.eye(3) if c > 0 else np.diag([1.0, -1.0, -1.0])
    else:
        vx = np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])
        r = np.eye(3) + vx + vx @ vx * (1.0 / (1.0 + c))
    m = np.eye(4)
    m[:3, :3] = r.T          # column-convention R -> row-convention R^T
    return m


def transform_aabb(aabb, m):
    lo, hi = np.asarray(aabb[:3]), np.asarray(aabb[3:])
    corners = np.array([[x, y, z] for x in (lo[0], hi[0]) for y in (lo[1], hi[1])
                        for z in (lo[2], hi[2])])
    w = corners @ m[:3, :3] + m[3, :3]
    return np.concatenate([w.min(0), w.max(0)])


# ---------------------------------------------------------------------------
# CSG node factories (CSG/CSGNode.cc:518-600 parameter conventions)
# ---------------------------------------------------------------------------
class Node:
    """A CSG node before flattening: leaf params or an operator with children."""

    def __init__(self, typecode, param=(0, 0, 0, 0, 0, 0), aabb=None, left=None,
                 right=None, complement=False, transform=None, planes=None, subs=None):
        self.typecode = typecode
        self.subs = subs                      # leaf Nodes of a list node (CSG_CONTIGUOUS..OVERLAP)
        self.param = list(param)
        self.aabb = None if aabb is None else np.asarray(aabb, dtype=np.float64)
        self.left, self.right = left, right
        self.complement = complement
        self.transform = transform            # 4x4 row-convention or None
        self.planes = planes                  # (n,4) for convexpolyhedron

    # -- operators --------------------------------------------------------
    def __or__(self, o):
        return Node(CSG_UNION, left=self, right=o)

    def __and__(self, o):
        return Node(CSG_INTERSECTION, left=self, right=o)

    def __sub__(self, o):
        return Node(CSG_DIFFERENCE, left=self, right=o)

    def moved(self, m):
        """Return self with transform m applied after any existing one."""
        if self.subs is not None:
            return Node(self.typecode, subs=[n.moved(m) for n in self.subs])
        if self.typecode < CSG_SPHERE and self.typecode != CSG_ZERO:
            return Node(self.typecode, left=self.left.moved(m), right=self.right.moved(m),
                        complement=self.complement)
        t = m if self.transform is None else self.transform @ m
        n = Node(self.typecode, self.param, self.aabb, complement=self.complement,
                 transform=t, planes=self.planes)
        return n

    def height(self):
        if self.left is None:
            return 0
        return 1 + max(self.left.height(), self.right.height())


def sphere(r, cx=0.0, cy=0.0, cz=0.0):
    return Node(CSG_SPHERE, (cx, cy, cz, r, 0, 0), (cx - r, cy - r, cz - r, cx + r, cy + r, cz + r))


def zsphere(r, z1, z2):
    assert z2 > z1
    return Node(CSG_ZSPHERE, (0, 0, 0, r, z1, z2), (-r, -r, z1, r, r, z2))


def cylinder(r, z1, z2):
    assert z2 > z1
    return Node(CSG_CYLINDER, (0, 0, 0, r, z1, z2), (-r, -r, z1, r, r, z2))


def box3(fx, fy=None, fz=None):
    fy = fx if fy is None else fy
    fz = fx if fz is None else fz
    return Node(CSG_BOX3, (fx, fy, fz, 0, 0, 0), (-fx / 2, -fy / 2, -fz / 2, fx / 2, fy / 2, fz / 2))


def cone(r1, z1, r2, z2):
    assert z2 > z1
    rm = max(r1, r2)
    return Node(CSG_CONE, (r1, z1, r2, z2, 0, 0), (-rm, -rm, z1, rm, rm, z2))


def convexpolyhedron(planes, aabb):
    planes = np.asarray(planes, dtype=np.float64)
    return Node(CSG_CONVEXPOLYHEDRON, (0, 0, 0, 0, 0, 0), aabb, planes=planes)


def hyperboloid(r0, zf, z1, z2):
    """One-sheet hyperboloid x^2+y^2 = r0^2((z/zf)^2+1) cut by the planes z1 < z2."""
    R = float(np.sqrt(r0 * r0 * (max(z1 * z1, z2 * z2) / (zf * zf) + 1.0)))
    return Node(CSG_HYPERBOLOID, (r0, zf, z1, z2, 0, 0), (-R, -R, z1, R, R, z2))


def phicut(phi0_pi, phi1_pi):
    """Unbounded phi wedge between phi0 and phi1 (units of pi), csg_intersect_leaf_phicut.h."""
    c0, s0 = np.cos(phi0_pi * np.pi), np.sin(phi0_pi * np.pi)
    c1, s1 = np.cos(phi1_pi * np.pi), np.sin(phi1_pi * np.pi)
    big = 1e6
    return Node(CSG_PHICUT, (c0, s0, c1, s1, 0, 0), (-big, -big, -big, big, big, big))


def halfspace(nx, ny, nz, w):
    """Halfspace p.n < w behind the plane p.n = w, normal normalised (sysrap/sn.h:3417-3426,
    CSG/csg_intersect_leaf_halfspace.h). Unbounded: its node box is zero as in
    CSGNode::HalfSpace (CSG/CSGNode.cc:669-676), so a prim needs a bounded operand to
    intersect it with."""
    v = np.array([nx, ny, nz], dtype=np.float64)
    v = v / np.linalg.norm(v)
    return Node(CSG_HALFSPACE, (v[0], v[1], v[2], w, 0, 0), None)


def complement(n):
    if n.subs is not None:
        raise ValueError('list nodes cannot be complemented')
    n2 = Node(n.typecode, n.param, n.aabb, n.left, n.right, not n.complement, n.transform, n.planes)
    return n2


def _list(typecode, leaves):
    leaves = list(leaves)
    if not leaves or any(l.left is not None or l.subs is not None for l in leaves):
        raise ValueError('a list node takes one or more leaf nodes')
    return Node(typecode, subs=leaves)


def contiguous(*leaves):
    """Union of touching or overlapping leaves as one list node (CSG_CONTIGUOUS, <= 8 leaves)."""
    return _list(CSG_CONTIGUOUS, leaves)


def discontiguous(*leaves):
    """Union of leaves that do not touch (CSG_DISCONTIGUOUS)."""
    return _list(CSG_DISCONTIGUOUS, leaves)


def overlap(*leaves):
    """Intersection of leaves (CSG_OVERLAP)."""
    return _list(CSG_OVERLAP, leaves)


# ---------------------------------------------------------------------------
def _positivize(n: Node, neg=False) -> Node:
    """Rewrite differences as intersections with complements (sysrap/sn.h:2694-2760):
    A - B = A & !B, and push complements down with De Morgan."""
    if n.left is None:
        if neg:
            return complement(n)
        return n
    t = n.typecode
    if t == CSG_DIFFERENCE:
        l = _positivize(n.left, neg)
        r = _positivize(n.right, not neg)
        op = CSG_UNION if neg else CSG_INTERSECTION
    else:
        l = _positivize(n.left, neg)
        r = _positivize(n.right, neg)
        op = t
        if neg:
            op = CSG_UNION if t == CSG_INTERSECTION else CSG_INTERSECTION
    return Node(op, left=l, right=r)


MAX_TREE_HEIGHT = 7      # complete binary tree of 255 nodes, the largest a prim may hold


def _operands(n: Node, op):
    if n.left is not None and n.typecode == op:
        return _operands(n.left, op) + _operands(n.right, op)
    return [n]


def _balanced(n: Node) -> Node:
    """Rebuild each chain of one operator as a balanced tree. Applied to positivised
    trees only, where every operator is a union or an intersection (associative and
    commutative), so the solid is unchanged; a chain of k operands drops from height
    k-1 to ceil(log2 k). Long subtraction chains (a plate minus many holes) need it."""
    if n.left is None:
        return n
    items = [_balanced(x) for x in _operands(n, n.typecode)]

    def build(xs):
        if len(xs) == 1:
            return xs[0]
        h = len(xs) // 2
        return Node(n.typecode, left=build(xs[:h]), right=build(xs[h:]))
    return build(items)


def _leaf_aabb(n: Node):
    return n.aabb if n.transform is None else transform_aabb(n.aabb, n.transform)


def _aabb_of(n: Node):
    """Conservative world AABB of a positivised tree; None means unbounded."""
    if n.subs is not None:
        bs = [_leaf_aabb(x) for x in n.subs if not x.complement]
        if n.typecode == CSG_OVERLAP:
            return np.concatenate([np.max([b[:3] for b in bs], axis=0), np.min([b[3:] for b in bs], axis=0)])
        if len(bs) < len(n.subs):
            return None
        return np.concatenate([np.min([b[:3] for b in bs], axis=0), np.max([b[3:] for b in bs], axis=0)])
    if n.left is None:
        if n.complement or n.aabb is None:
            return None
        a = n.aabb
        if n.transform is not None:
            a = transform_aabb(a, n.transform)
        return a
    la, ra = _aabb_of(n.left), _aabb_of(n.right)
    if n.typecode == CSG_UNION:
        if la is None or ra is None:
            return None
        return np.concatenate([np.minimum(la[:3], ra[:3]), np.maximum(la[3:], ra[3:])])
    # intersection
    if la is None:
        return ra
    if ra is None:
        return la
    return np.concatenate([np.maximum(la[:3], ra[:3]), np.minimum(la[3:], ra[3:])])


class Foundry:
    """Accumulates solids/prims/nodes/transforms/instances + materials/surfaces."""

    def __init__(self):
        self.nodes = []        # list of float32[16]
        self.prims = []
        self.solids = []
        self.tran = []         # 4x4 float64 (row convention), 1-based via gtransformIdx
        self.plan = []
        self.inst = []
        self.materials = {}    # name -> dict of arrays over domain
        self.mat_order = []
        self.surfaces = {}     # name -> (payload (nw,4), osn0)
        self.sur_order = []
        self.boundaries = []   # (omat, osur, isur, imat) names
        self._cur_solid = None

    # -- materials / surfaces ---------------------------------------------
    def add_material(self, name, rindex, absorption_length=1e6, scattering_length=1e6,
                     reemission_prob=0.0, group_velocity=None):
        """Material rows over the wavelength domain; each property is a constant, a
        function of wavelength (nm) or an array over the domain. Without group_velocity
        it is derived from rindex as c/(n - lambda dn/dlambda)."""
        wl = domain()

        def arr(v):
            if callable(v):
                return np.asarray(v(wl), dtype=np.float64)
            if np.ndim(v):
                a = np.asarray(v, dtype=np.float64)
                assert a.shape == (NUM_WL,), a.shape
                return a
            return np.full(NUM_WL, float(v))

        n = arr(rindex)
        if group_velocity is None:
            dn = np.gradient(n, wl)
            vg = C_LIGHT / (n - wl * dn)
        else:
            vg = arr(group_velocity)
        self.materials[name] = dict(rindex=n, abs=arr(absorption_length),
                                    ray=arr(scattering_length), reemit=arr(reemission_prob),
                                    vg=vg)
        self.mat_order.append(name)

    def add_surface(self, name, detect=0.0, absorb=0.0, specular=0.0, diffuse=0.0, osn0='X',
                    finish=0, model_value_percent=0):
        """Surface payload as U4SurfaceArray::addSurface: sensor -> (eff, 1-eff, 0, 0);
        polished -> (0, 1-R, R, 0); ground -> (0, 1-R, 0, R). Each component is a constant
        or an array over the wavelength domain. finish and model_value_percent fill the
        optical row's z and w (sstandard::make_optical, sysrap/sstandard.h:262-380)."""
        p = np.zeros((NUM_WL, 4))
        for k, v in enumerate((detect, absorb, specular, diffuse)):
            p[:, k] = v
        assert np.all(np.abs(p.sum(axis=1) - 1.0) < 1e-9), name
        self.surfaces[name] = (p, osn0, int(finish), int(model_value_percent))
        self.sur_order.append(name)

    def add_sensor_surface(self, name, efficiency, osn0='X'):
        self.add_surface(name, detect=efficiency, absorb=1.0 - efficiency, osn0=osn0)

    def add_implicit_absorber(self, name):
        """U4SurfaceArray::addImplicit: perfect absorber for RINDEX-less materials."""
        self.add_surface(name, absorb=1.0, osn0='X', finish=1, model_value_percent=100)

    def boundary(self, omat, osur, isur, imat):
        key = (omat, osur, isur, imat)
        if key not in self.boundaries:
            self.boundaries.append(key)
        return self.boundaries.index(key)

    # -- geometry -----------------------------------------------------------
    def _add_transform(self, m):
        self.tran.append(np.asarray(m, dtype=np.float64))
        return len(self.tran)          # 1-based

    def begin_solid(self, label):
        self._cur_solid = dict(label=label, primOffset=len(self.prims), numPrim=0)

    def add_prim(self, tree: Node, boundary: int, mesh_idx=0):
        assert self._cur_solid is not None
        t = _positivize(tree)
        if t.height() > MAX_TREE_HEIGHT:
            t = _balanced(t)
        h = t.height()
        if h > MAX_TREE_HEIGHT:
            raise ValueError(f'CSG tree of height {h} exceeds {MAX_TREE_HEIGHT} after balancing')
        num = (1 << (h + 1)) - 1
        flat = [None] * num

        def place(n, i):                      # i is 1-based level order index
            flat[i - 1] = n
            if n.left is not None:
                place(n.left, 2 * i)
                place(n.right, 2 * i + 1)

        place(t, 1)
        # list nodes keep their leaves after the tree nodes: subOffset counts from the
        # prim's first node (CSGNode subNum/subOffset, CSG/CSGNode.h:187-211)
        extra = []
        sub_at = {}
        for i, n in enumerate(flat):
            if n is not None and n.subs is not None:
                sub_at[i] = num + len(extra)
                extra.extend(n.subs)
        node_offset = len(self.nodes)
        for i, n in enumerate(flat + extra):
            q = np.zeros(16, dtype=np.float32)
            u = q.view(np.uint32)
            if n is None:
                u[14] = CSG_ZERO
            elif n.subs is not None:
                u[14] = n.typecode
                u[0] = len(n.subs)
                u[1] = sub_at[i]
                aabb = _aabb_of(n)
                q[8:12] = aabb[0:4]
                q[12:14] = aabb[4:6]
            elif n.left is not None:
                u[14] = n.typecode
                if i == 0:
                    u[0] = num                  # subNum on the root operator
                    u[1] = 0                    # subOffset
            else:
                q[0:4] = n.param[0:4]
                q[4:6] = n.param[4:6]
                if n.aabb is not None:      # unbounded leaves (halfspace) keep a zero box
                    aabb = n.aabb if n.transform is None else transform_aabb(n.aabb, n.transform)
                    q[8:12] = aabb[0:4]
                    q[12:14] = aabb[4:6]
                u[14] = n.typecode
                tidx = 0
                if n.transform is not None and not np.allclose(n.transform, np.eye(4)):
                    tidx = self._add_transform(n.transform)
                if n.typecode == CSG_CONVEXPOLYHEDRON:
                    u[0] = len(self.plan)
                    u[1] = len(n.planes)
                    for pl in n.planes:
                        self.plan.append(np.asarray(pl, dtype=np.float64))
                u[15] = (tidx & 0x7fffffff) | (0x80000000 if n.complement else 0)
            if n is not None:
                u[6] = boundary
                u[7] = i
            self.nodes.append(q)

        aabb = _aabb_of(t)
        assert aabb is not None, "prim must be bounded"
        p = np.zeros(16, dtype=np.float32)
        pi = p.view(np.int32)
        pu = p.view(np.uint32)
        pi[0] = num + len(extra)
        pi[1] = node_offset
        pi[2] = 0
        pi[3] = 0
        local = self._cur_solid['numPrim']
        pu[4] = local                      # sbtIndexOffset
        pu[5] = mesh_idx
        pu[6] = len(self.solids)           # repeatIdx
        pu[7] = local                      # primIdx
        p[8:12] = aabb[0:4]
        p[12:14] = aabb[4:6]
        pu[15] = len(self.prims)           # globalPrimIdx
        self.prims.append(p)
        self._cur_solid['numPrim'] += 1
        return len(self.prims) - 1

    def end_solid(self):
        s = self._cur_solid
        self._cur_solid = None
        lo = np.array([np.inf] * 3)
        hi = -lo
        for p in self.prims[s['primOffset']:s['primOffset'] + s['numPrim']]:
            lo = np.minimum(lo, p[8:11])
            hi = np.maximum(hi, np.array([p[11], p[12], p[13]]))
        ce = np.concatenate([(lo + hi) / 2, [np.max(hi - lo) / 2]])
        rec = np.zeros(12, dtype=np.int32)
        lab = s['label'].encode()[:15]
        rec.view(np.uint8)[:len(lab)] = np.frombuffer(lab, dtype=np.uint8)
        rec[4] = s['numPrim']
        rec[5] = s['primOffset']
        rec[6] = 0
        rec[7] = 0
        rec[8:12] = ce.astype(np.float32).view(np.int32)
        self.solids.append(rec)
        return len(self.solids) - 1

    def add_instance(self, m, gas_idx, sensor_identifier=-1, sensor_index=-1):
        """qat4 with identity in .w column: ins_idx, gas_idx, sensor_identifier+1,
        sensor_index (sysrap/sqat4.h:386-393, CSGFoundry::addInstance)."""
        q = np.asarray(m, dtype=np.float64).astype(np.float32).copy()
        qi = q.view(np.int32)
        qi[0, 3] = len(self.inst)
        qi[1, 3] = gas_idx
        qi[2, 3] = sensor_identifier + 1
        qi[3, 3] = sensor_index
        self.inst.append(q)
        return len(self.inst) - 1

    # -- export ---------------------------------------------------------------
    def arrays(self):
        itra = []
        tran = []
        for m in self.tran:
            tran.append(m.astype(np.float32))
            itra.append(np.linalg.inv(m).astype(np.float32))
        g = dict(
            solid=np.array(self.solids, dtype=np.int32).reshape(-1, 3, 4),
            prim=np.array(self.prims, dtype=np.float32).reshape(-1, 4, 4),
            node=np.array(self.nodes, dtype=np.float32).reshape(-1, 4, 4),
            plan=np.array(self.plan, dtype=np.float32).reshape(-1, 4),
            tran=np.array(tran, dtype=np.float32).reshape(-1, 4, 4),
            itra=np.array(itra, dtype=np.float32).reshape(-1, 4, 4),
            inst=np.array(self.inst, dtype=np.float32).reshape(-1, 4, 4),
        )
        return g

    def physics(self):
        nb = len(self.boundaries)
        bnd = np.full((nb, 4, 2, NUM_WL, 4), -1.0, dtype=np.float64)
        optical = np.zeros((nb * 4, 4), dtype=np.uint32)
        for b, (om, os_, is_, im) in enumerate(self.boundaries):
            for j, name in enumerate((om, os_, is_, im)):
                line = 4 * b + j
                if j in (0, 3):
                    m = self.materials[name]
                    bnd[b, j, 0, :, 0] = m['rindex']
                    bnd[b, j, 0, :, 1] = m['abs']
                    bnd[b, j, 0, :, 2] = m['ray']
                    bnd[b, j, 0, :, 3] = m['reemit']
                    bnd[b, j, 1, :, :] = 0.0
                    bnd[b, j, 1, :, 0] = m['vg']
                    optical[line] = (self.mat_order.index(name) + 1, 0, 0, 0)
                else:
                    if name is None:
                        optical[line] = (0, SMATSUR_NOSURFACE, 0, 0)
                        continue
                    payload, osn0, finish, mvp = self.surfaces[name]
                    bnd[b, j, 0] = payload
                    bnd[b, j, 1] = 0.0
                    ems = {'-': SMATSUR_NOSURFACE, '#': SMATSUR_SENSOR_A}.get(osn0, SMATSUR_SURFACE)
                    optical[line] = (self.sur_order.index(name) + 1, ems, finish, mvp)
        return dict(bnd=bnd.astype(np.float32), optical=optical,
                    wl_min=WL_MIN, wl_step=WL_STEP)


# ---------------------------------------------------------------------------
def scint_icdf(spectrum, hd_factor=20, num_bins=4096):
    """ICDF of a scintillation spectrum given as intensity(wavelength_nm), sampled as
    U4Scint::CreateGeant4InterpolatedInverseCDF (u4/U4Scint.h:405-480): the integral is
    taken over photon energy; rows are u in [0,1), [0,1/hd), [1-1/hd, 1)."""
    e = np.linspace(HC_EVNM / 820.0, HC_EVNM / 60.0, 20001)
    val = spectrum(HC_EVNM / e)
    cum = np.concatenate([[0.0], np.cumsum(0.5 * (val[1:] + val[:-1]) * np.diff(e))])
    mx = cum[-1]
    j = np.arange(num_bins, dtype=np.float64)
    edge = 1.0 / hd_factor
    u_all = j / num_bins
    u_lhs = j / (hd_factor * num_bins)
    u_rhs = 1.0 - edge + j / (hd_factor * num_bins)

    def wl(u):
        en = np.interp(u * mx, cum, e)
        return HC_EVNM / en

    icdf = np.stack([wl(u_all), wl(u_lhs), wl(u_rhs)]).astype(np.float32)
    return icdf
