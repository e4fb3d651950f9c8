"""GDML geometry front-end: a GDML file in, CSGFoundry arrays and SSim tables out.

The reference reaches CSGFoundry from GDML through Geant4: G4GDMLParser builds the
G4 volume tree, U4Tree turns it into an stree (u4/U4Tree.h:790-900), U4Solid converts each
G4VSolid into an sn CSG tree (u4/U4Solid.h:335-1112), U4SurfaceArray/U4Material standardise
surfaces and materials over the wavelength domain (u4/U4SurfaceArray.h:147-283,
u4/U4Material.cc:622-700), and CSGImport serialises the trees (CSG/CSGImport.cc:82-600).
Geant4 is not part of this build, so this module reads the GDML itself and applies the
same conversion rules:

* names lose their "0x..." pointer suffix, as G4GDMLRead::StripName does;
* placements and boolean displacements use the GDML rotation inverted, as G4GDMLRead
  does (rotation = Rz Ry Rx of the given angles);
* solids: box -> box3; orb -> sphere; sphere -> sphere or zsphere (theta cut), minus the
  inner sphere; ellipsoid -> sphere or zsphere of radius c scaled by (a/c, b/c, 1), a
  one-sided z cut widened by 0.1 mm on the uncut side; tube -> cylinder minus the inner
  cylinder lengthened by 1% each end; cone -> cone minus inner cone; union, subtraction,
  intersection with the second solid displaced. Phi segments of tubes and cones are
  ignored, as U4Solid ignores them (a warning is recorded); a phi-segmented sphere is an
  error, as in U4Solid::init_Sphere_.
* extensions beyond U4Solid (recorded in GdmlGeometry.notes): trd and trap become convex
  polyhedra; polycone becomes the union of its z segments (cones/cylinders) minus the
  union of its inner segments, without U4Polycone's z nudging;
* boundaries (omat, osur, isur, imat) per placed volume as U4TreeBorder: osur is the
  border or skin surface found for (parent -> volume), isur for (volume -> parent) when
  the volume's material has RINDEX, looked up like G4OpBoundaryProcess
  (U4Surface::Find, u4/U4Surface.h:422-443); a RINDEX -> no-RINDEX step with no surface
  gets an implicit perfect absorber (u4/U4TreeBorder.h:100-260);
* surfaces are ordered border surfaces, skin surfaces, then implicits; materials in the
  order a postorder walk first meets them (U4Tree::initMaterials_r);
* material rows RINDEX, ABSLENGTH, RAYLEIGH, REEMISSIONPROB, GROUPVEL sampled at the
  domain energies hc/lambda by linear interpolation in energy clamped at the ends
  (G4PhysicsVector::Value), defaults 1, 1e12, 1e12, 0, c (sysrap/sproplist.h:31-40);
  a material without GROUPVEL gets one derived from RINDEX (c / (n + dn/dlnE) at the
  energy midpoints, the Geant4 G4MaterialPropertiesTable::CalculateGROUPVEL rule);
* surface payloads (detect, absorb, reflect_specular, reflect_diffuse) from EFFICIENCY
  and REFLECTIVITY as U4SurfaceArray::addSurface; optical rows (index+1, ems from the
  optical surface name's first character, finish, 100*value) as sstandard::make_optical;
* placed volumes become prims of the global solid in preorder; subtrees repeated at least
  freq_cut times (stree FREQ_CUT, default 500, sysrap/stree.h:293-295) become instanced
  solids, one prim per volume of the subtree in the subtree root's frame;
* the first material with a scintillation spectrum (FASTCOMPONENT or
  SCINTILLATIONCOMPONENT1) supplies the 3-row HD20 ICDF (QSim::UploadComponents).

Trees deeper than a prim holds are rebalanced (csg._balanced), which the reference does
not do; it applies sn::uncoincide z nudges, which this module does not.
"""
from __future__ import annotations

import ast
import math
import operator
import re
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import csg

HC_EVNM = 1239.84198433200208455673            # sysrap/sdomain.h:25
C_LIGHT = 299.792458                            # mm/ns
MAX_SUBS_CONTIGUOUS = 8

# CLHEP units in Geant4's internal system (mm, ns, MeV, rad)
UNITS = {
    'mm': 1.0, 'millimeter': 1.0, 'cm': 10.0, 'centimeter': 10.0, 'm': 1000.0, 'meter': 1000.0,
    'km': 1e6, 'um': 1e-3, 'micrometer': 1e-3, 'nm': 1e-6, 'nanometer': 1e-6, 'angstrom': 1e-7,
    'fermi': 1e-12, 'pc': 3.0856775807e19, 'parsec': 3.0856775807e19,
    'rad': 1.0, 'radian': 1.0, 'mrad': 1e-3, 'milliradian': 1e-3, 'deg': math.pi / 180.0,
    'degree': math.pi / 180.0, 'sr': 1.0, 'steradian': 1.0,
    'ns': 1.0, 'nanosecond': 1.0, 's': 1e9, 'second': 1e9, 'ms': 1e6, 'millisecond': 1e6,
    'us': 1e3, 'microsecond': 1e3, 'ps': 1e-3, 'picosecond': 1e-3,
    'MeV': 1.0, 'megaelectronvolt': 1.0, 'eV': 1e-6, 'electronvolt': 1e-6, 'keV': 1e-3,
    'kiloelectronvolt': 1e-3, 'GeV': 1e3, 'gigaelectronvolt': 1e3, 'TeV': 1e6, 'PeV': 1e9,
    'g': 6.241509074e21, 'gram': 6.241509074e21, 'kg': 6.241509074e24, 'mg': 6.241509074e18,
    'mole': 1.0, 'kelvin': 1.0, 'K': 1.0, 'perCent': 0.01, 'percent': 0.01, 'perThousand': 1e-3,
    'cm2': 100.0, 'cm3': 1000.0, 'm2': 1e6, 'm3': 1e9, 'mm2': 1.0, 'mm3': 1.0,
    'pascal': 6.241509074e-10, 'bar': 6.241509074e-5, 'atmosphere': 6.32420e-5,
    'pi': math.pi, 'twopi': 2 * math.pi, 'halfpi': math.pi / 2, 'e': math.e,
}
UNITS['g/cm3'] = UNITS['g'] / UNITS['cm3']
UNITS['g/mole'] = UNITS['g'] / UNITS['mole']

_FUNCS = {
    'sin': math.sin, 'cos': math.cos, 'tan': math.tan, 'asin': math.asin, 'acos': math.acos,
    'atan': math.atan, 'atan2': math.atan2, 'sinh': math.sinh, 'cosh': math.cosh,
    'tanh': math.tanh, 'exp': math.exp, 'log': math.log, 'log10': math.log10,
    'sqrt': math.sqrt, 'pow': math.pow, 'abs': abs, 'fabs': abs, 'min': min, 'max': max,
    'floor': math.floor, 'ceil': math.ceil,
}
_BIN = {ast.Add: operator.add, ast.Sub: operator.sub, ast.Mult: operator.mul,
        ast.Div: operator.truediv, ast.Pow: operator.pow, ast.Mod: math.fmod}
_UN = {ast.UAdd: operator.pos, ast.USub: operator.neg}


def strip_name(name: Optional[str]) -> Optional[str]:
    """G4GDMLRead::StripName: drop everything from the first "0x"."""
    if name is None:
        return None
    k = name.find('0x')
    return name if k < 0 else name[:k]


class GdmlError(ValueError):
    pass


class Evaluator:
    """Arithmetic over GDML constants, variables, quantities and CLHEP units, the
    expression language of G4GDMLEvaluator (numbers, + - * / ^ **, unary signs,
    parentheses, names and the usual math functions). Nothing else is evaluated."""

    def __init__(self):
        self.vars: Dict[str, float] = dict(UNITS)
        self.loop: Dict[str, float] = {}

    def __call__(self, expr, default=None) -> float:
        if expr is None:
            if default is None:
                raise GdmlError('missing value')
            return float(default)
        if isinstance(expr, (int, float)):
            return float(expr)
        text = str(expr).strip().replace('^', '**')
        if not text:
            raise GdmlError('empty expression')
        try:
            tree = ast.parse(text, mode='eval')
        except SyntaxError as e:
            raise GdmlError(f'cannot parse expression {expr!r}') from e
        return float(self._eval(tree.body, expr))

    def _eval(self, n, src):
        if isinstance(n, ast.Constant) and isinstance(n.value, (int, float)):
            return n.value
        if isinstance(n, ast.BinOp) and type(n.op) in _BIN:
            return _BIN[type(n.op)](self._eval(n.left, src), self._eval(n.right, src))
        if isinstance(n, ast.UnaryOp) and type(n.op) in _UN:
            return _UN[type(n.op)](self._eval(n.operand, src))
        if isinstance(n, ast.Name):
            if n.id in self.loop:
                return self.loop[n.id]
            if n.id in self.vars:
                return self.vars[n.id]
            stripped = strip_name(n.id)
            if stripped in self.vars:
                return self.vars[stripped]
            raise GdmlError(f'unknown name {n.id!r} in {src!r}')
        if isinstance(n, ast.Call) and isinstance(n.func, ast.Name) and n.func.id in _FUNCS and not n.keywords:
            return _FUNCS[n.func.id](*[self._eval(a, src) for a in n.args])
        raise GdmlError(f'unsupported expression {src!r}')

    def name(self, s: Optional[str]) -> Optional[str]:
        """A name with the loop variables of the current iteration written in, {i} or [i].
        References resolve on full names; the pointer suffix is stripped only from the
        names that reach the output, as G4GDMLParser strips its stores after reading."""
        if s is None:
            return None
        for k, v in self.loop.items():
            iv = str(int(v)) if float(v).is_integer() else str(v)
            s = s.replace('{' + k + '}', iv).replace('[' + k + ']', iv)
        return s


# ------------------------------------------------------------------ rotations
def _rx(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])


def _ry(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])


def _rz(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])


def placement(pos, angles, scale=(1.0, 1.0, 1.0)):
    """Row-convention 4x4 object transform of a GDML placement or boolean displacement:
    G4Transform3D(R^-1, pos) with R = Rz Ry Rx (G4GDMLRead GetRotationMatrix, then
    inverse), times the scale. Row convention p' = p M, so M's 3x3 block is R itself."""
    R = _rz(angles[2]) @ _ry(angles[1]) @ _rx(angles[0])
    m = np.eye(4)
    m[:3, :3] = np.diag(scale) @ R
    m[3, :3] = pos
    return m


# ------------------------------------------------------------------ result
@dataclass
class GdmlGeometry:
    geometry: Dict[str, np.ndarray]
    physics: Dict[str, object]
    materials: List[str]
    surfaces: List[str]
    boundaries: List[str]                  # "omat/osur/isur/imat"
    prim_volumes: List[str]                # physical volume name of each prim
    prim_boundary: List[int]
    num_factor: int = 0
    notes: List[str] = field(default_factory=list)

    def boundary_index(self, spec: str) -> int:
        return self.boundaries.index(spec)


@dataclass
class _Volume:
    name: str
    material: str
    solid: str
    children: list                         # [(pv name, lv name, local 4x4, copyno)]


@dataclass
class _PlacedNode:
    index: int
    pv: str
    lv: str
    m2w: np.ndarray
    parent: int
    depth: int
    copyno: int
    children: list = field(default_factory=list)
    boundary: int = -1
    digest: str = ''


class GdmlReader:
    """Reads the define, materials, solids, structure and setup sections of a GDML file."""

    def __init__(self, path):
        self.path = str(path)
        self.ev = Evaluator()
        self.positions: Dict[str, np.ndarray] = {}
        self.rotations: Dict[str, np.ndarray] = {}
        self.scales: Dict[str, np.ndarray] = {}
        self.matrices: Dict[str, np.ndarray] = {}
        self.materials: Dict[str, Dict[str, np.ndarray]] = {}
        self.solids: Dict[str, ET.Element] = {}
        self.solid_loop: Dict[str, Dict[str, float]] = {}
        self.optical_surfaces: Dict[str, dict] = {}
        self.volumes: Dict[str, _Volume] = {}
        self.border: List[tuple] = []            # (surface name, pv1, pv2, optical surface)
        self.skin: List[tuple] = []              # (surface name, lv, optical surface)
        self.world: Optional[str] = None
        self.notes: List[str] = []
        try:
            root = ET.parse(self.path).getroot()
        except ET.ParseError as e:
            raise GdmlError(f'{self.path}: {e}') from e
        sections = {child.tag: child for child in root}
        for tag in ('define', 'materials', 'solids', 'structure', 'setup'):
            if tag in sections:
                getattr(self, '_read_' + tag)(sections[tag])
        if self.world is None:
            raise GdmlError(f'{self.path}: no <setup><world ref=.../></setup>')

    # -- loops: yield each child with the loop variables of its iteration set
    def _children(self, elem):
        for c in elem:
            if c.tag == 'loop':
                var = c.get('for')
                lo, hi = self.ev(c.get('from')), self.ev(c.get('to'))
                step = self.ev(c.get('step'), 1.0)
                if step == 0 or var is None:
                    raise GdmlError('loop with zero step or no variable')
                saved = self.ev.loop.get(var)
                v = lo
                while (step > 0 and v <= hi + 1e-9) or (step < 0 and v >= hi - 1e-9):
                    self.ev.loop[var] = v
                    yield from self._children(c)
                    v += step
                if saved is None:
                    self.ev.loop.pop(var, None)
                else:
                    self.ev.loop[var] = saved
            elif isinstance(c.tag, str):
                yield c

    def _vec3(self, e, unit_attr='unit', default_unit='mm'):
        u = self.ev(e.get(unit_attr, default_unit))
        return np.array([self.ev(e.get(k), 0.0) * u for k in ('x', 'y', 'z')])

    def _read_define(self, d):
        for e in self._children(d):
            name = self.ev.name(e.get('name'))
            if e.tag in ('constant', 'variable'):
                self.ev.vars[name] = self.ev(e.get('value'))
            elif e.tag == 'quantity':
                self.ev.vars[name] = self.ev(e.get('value')) * self.ev(e.get('unit', '1'))
            elif e.tag == 'expression':
                self.ev.vars[name] = self.ev(e.text)
            elif e.tag == 'position':
                self.positions[name] = self._vec3(e)
            elif e.tag == 'rotation':
                self.rotations[name] = self._vec3(e, default_unit='rad')
            elif e.tag == 'scale':
                self.scales[name] = np.array([self.ev(e.get(k), 1.0) for k in ('x', 'y', 'z')])
            elif e.tag == 'matrix':
                coldim = int(self.ev(e.get('coldim'), 1))
                vals = [self.ev(t) for t in (e.get('values') or e.get('value') or '').split()]
                if coldim < 1 or len(vals) % coldim:
                    raise GdmlError(f'matrix {name}: {len(vals)} values do not fill coldim {coldim}')
                self.matrices[name] = np.asarray(vals, dtype=np.float64).reshape(-1, coldim)

    def _props(self, e):
        out = {}
        for p in e.findall('property'):
            ref = p.get('ref')
            if ref is None and p.get('value') is not None:       # inline constant property
                out[p.get('name')] = np.array([[self.ev(p.get('value'))]])
                continue
            if ref not in self.matrices:
                raise GdmlError(f'property {p.get("name")} refers to unknown matrix {ref}')
            out[p.get('name')] = self.matrices[ref]
        return out

    def _read_materials(self, m):
        for e in self._children(m):
            if e.tag == 'material':
                self.materials[self.ev.name(e.get('name'))] = self._props(e)

    def _read_solids(self, s):
        for e in self._children(s):
            name = self.ev.name(e.get('name'))
            if e.tag == 'opticalsurface':
                self.optical_surfaces[name] = dict(
                    name=name, model=e.get('model', 'glisur'), finish=e.get('finish', 'polished'),
                    type=e.get('type', 'dielectric_dielectric'), value=self.ev(e.get('value'), 1.0),
                    props=self._props(e))
            else:
                self.solids[name] = e
                self.solid_loop[name] = dict(self.ev.loop)

    def _transform_of(self, e, prefix=''):
        pos, rot, sc = np.zeros(3), np.zeros(3), np.ones(3)
        for c in e:
            tag = c.tag
            if tag == prefix + 'position':
                pos = self._vec3(c)
            elif tag == prefix + 'positionref':
                pos = self.positions[c.get('ref')]
            elif tag == prefix + 'rotation':
                rot = self._vec3(c, default_unit='rad')
            elif tag == prefix + 'rotationref':
                rot = self.rotations[c.get('ref')]
            elif tag == prefix + 'scale':
                sc = np.array([self.ev(c.get(k), 1.0) for k in ('x', 'y', 'z')])
            elif tag == prefix + 'scaleref':
                sc = self.scales[c.get('ref')]
        if np.any(sc <= 0):
            raise GdmlError('reflections (negative scale) are not supported')
        return placement(pos, rot, sc)

    def _read_structure(self, st):
        for e in self._children(st):
            if e.tag == 'volume':
                name = self.ev.name(e.get('name'))
                mat = e.find('materialref')
                sol = e.find('solidref')
                if mat is None or sol is None:
                    raise GdmlError(f'volume {name} without materialref or solidref')
                kids = []
                for pv in self._children(e):
                    if pv.tag in ('replicavol', 'paramvol', 'divisionvol'):
                        raise GdmlError(f'volume {name}: {pv.tag} is not supported')
                    if pv.tag != 'physvol':
                        continue
                    ref = pv.find('volumeref')
                    if ref is None:
                        raise GdmlError(f'physvol in {name} without volumeref (file/assembly refs are not supported)')
                    lv = self.ev.name(ref.get('ref'))
                    pvname = self.ev.name(pv.get('name')) or f'{lv}_PV'
                    copyno = int(self.ev(pv.get('copynumber'), 0))
                    kids.append((pvname, lv, self._transform_of(pv), copyno))
                self.volumes[name] = _Volume(name, self.ev.name(mat.get('ref')), self.ev.name(sol.get('ref')), kids)
            elif e.tag == 'assembly':
                raise GdmlError('assembly volumes are not supported')
            elif e.tag == 'bordersurface':
                refs = [self.ev.name(r.get('ref')) for r in e.findall('physvolref')]
                if len(refs) != 2:
                    raise GdmlError('bordersurface needs two physvolref')
                self.border.append((self.ev.name(e.get('name')), refs[0], refs[1],
                                    self.ev.name(e.get('surfaceproperty'))))
            elif e.tag == 'skinsurface':
                ref = e.find('volumeref')
                self.skin.append((self.ev.name(e.get('name')), self.ev.name(ref.get('ref')),
                                  self.ev.name(e.get('surfaceproperty'))))

    def _read_setup(self, s):
        w = s.find('world')
        if w is not None:
            self.world = self.ev.name(w.get('ref'))

    # -- solids -> CSG trees
    def solid(self, name) -> csg.Node:
        if name not in self.solids:
            raise GdmlError(f'unknown solid {name}')
        saved = self.ev.loop
        self.ev.loop = dict(self.solid_loop[name])
        try:
            return self._solid(self.solids[name])
        finally:
            self.ev.loop = saved

    def _len(self, e, k, default=0.0):
        return self.ev(e.get(k), default) * self.ev(e.get('lunit', 'mm'))

    def _ang(self, e, k, default=0.0):
        return self.ev(e.get(k), default) * self.ev(e.get('aunit', 'rad'))

    def _phi_full(self, e):
        """A phi range within half of Geant4's angular tolerance (1e-9 rad) of 2 pi is
        the full circle, as G4Sphere/G4Tubs/G4Cons CheckPhiAngles snap it."""
        return self._ang(e, 'deltaphi', 2 * math.pi) >= 2 * math.pi - 0.5e-9

    def _solid(self, e) -> csg.Node:
        t = e.tag
        name = self.ev.name(e.get('name'))
        if t == 'box':
            return csg.box3(self._len(e, 'x'), self._len(e, 'y'), self._len(e, 'z'))
        if t == 'orb':
            return csg.sphere(self._len(e, 'r'))
        if t == 'sphere':
            if not self._phi_full(e):
                raise GdmlError(f'sphere {name}: phi segments are not supported (U4Solid::init_Sphere_)')
            outer = self._sphere_layer(e, 'rmax')
            inner = self._sphere_layer(e, 'rmin')
            return outer if inner is None else outer - inner
        if t == 'ellipsoid':
            return self._ellipsoid(e)
        if t == 'tube':
            if not self._phi_full(e):
                self.notes.append(f'tube {name}: phi segment ignored, as U4Solid::init_Tubs ignores it')
            hz = 0.5 * self._len(e, 'z')
            rmax, rmin = self._len(e, 'rmax'), self._len(e, 'rmin')
            outer = csg.cylinder(rmax, -hz, hz)
            if rmin <= 0.0:
                return outer
            dz = 0.01 * hz
            return outer - csg.cylinder(rmin, -(hz + dz), hz + dz)
        if t == 'cone':
            if not self._phi_full(e):
                self.notes.append(f'cone {name}: phi segment ignored, as U4Solid::init_Cons ignores it')
            hz = 0.5 * self._len(e, 'z')
            outer = csg.cone(self._len(e, 'rmax1'), -hz, self._len(e, 'rmax2'), hz)
            r1, r2 = self._len(e, 'rmin1'), self._len(e, 'rmin2')
            if r1 == 0.0 and r2 == 0.0:
                return outer
            return outer - csg.cone(r1, -hz, r2, hz)
        if t == 'polycone':
            return self._polycone(e, name)
        if t == 'trd':
            return self._trd(e, name)
        if t == 'trap':
            return self._trap(e, name)
        if t in ('union', 'subtraction', 'intersection'):
            first = self.solid(self.ev.name(e.find('first').get('ref')))
            second = self.solid(self.ev.name(e.find('second').get('ref')))
            m1 = self._transform_of(e, 'first')
            if not np.allclose(m1, np.eye(4)):
                first = first.moved(m1)
            m2 = self._transform_of(e)
            if not np.allclose(m2, np.eye(4)):
                second = second.moved(m2)
            return {'union': first | second, 'subtraction': first - second,
                    'intersection': first & second}[t]
        raise GdmlError(f'solid {name}: <{t}> is not supported')

    def _sphere_layer(self, e, key):
        r = self._len(e, key)
        if r == 0.0:
            return None
        st = self._ang(e, 'starttheta', 0.0)
        dt = self._ang(e, 'deltatheta', math.pi)
        if st + dt >= math.pi - 0.5e-9:            # G4Sphere snaps the theta range to pi
            dt = math.pi - st
        lt = st + dt
        if st > 0.0 or dt < math.pi:
            return csg.zsphere(r, r * math.cos(lt), r * math.cos(st))
        return csg.sphere(r)

    def _ellipsoid(self, e):
        a, b, c = self._len(e, 'ax'), self._len(e, 'by'), self._len(e, 'cz')
        z1, z2 = self._len(e, 'zcut1', 0.0), self._len(e, 'zcut2', 0.0)
        if z1 == 0.0 and z2 == 0.0:                 # G4Ellipsoid: no cuts
            z1, z2 = -c, c
        zmin = z1 if z1 > -c else -c
        zmax = z2 if z2 < c else c
        upper, lower = zmax < c, zmin > -c
        if not upper and not lower:
            n = csg.sphere(c)
        elif upper and lower:
            n = csg.zsphere(c, zmin, zmax)
        elif lower:
            n = csg.zsphere(c, zmin, zmax + 0.1)
        else:
            n = csg.zsphere(c, zmin - 0.1, zmax)
        return n.moved(csg.scale(a / c, b / c, 1.0))

    def _polycone(self, e, name):
        if not self._phi_full(e):
            self.notes.append(f'polycone {name}: phi segment ignored')
        lu = self.ev(e.get('lunit', 'mm'))
        planes = [(self.ev(z.get('rmin'), 0.0) * lu, self.ev(z.get('rmax')) * lu, self.ev(z.get('z')) * lu)
                  for z in e.findall('zplane')]
        if len(planes) < 2:
            raise GdmlError(f'polycone {name}: needs two zplanes')
        if planes[0][2] > planes[-1][2]:
            planes = planes[::-1]
        outer, inner = [], []
        for (ra0, rb0, z0), (ra1, rb1, z1) in zip(planes[:-1], planes[1:]):
            if z1 <= z0:
                continue
            outer.append(csg.cylinder(rb0, z0, z1) if rb0 == rb1 else csg.cone(rb0, z0, rb1, z1))
            if ra0 > 0.0 or ra1 > 0.0:
                inner.append(csg.cylinder(ra0, z0, z1) if ra0 == ra1 else csg.cone(ra0, z0, ra1, z1))
        self.notes.append(f'polycone {name}: union of {len(outer)} segments (no z nudging)')

        def union(xs):
            if len(xs) == 1:
                return xs[0]
            h = len(xs) // 2
            return union(xs[:h]) | union(xs[h:])
        n = union(outer)
        return n - union(inner) if inner else n

    @staticmethod
    def _hull(name, verts):
        """Outward planes (nx, ny, nz, d) of the convex hull of 8 hexahedron vertices in
        G4 order (-z face: -y-x, -y+x, +y-x, +y+x; then the +z face)."""
        v = np.asarray(verts, dtype=np.float64)
        faces = [(0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (2, 6, 7, 3), (0, 4, 6, 2), (1, 3, 7, 5)]
        c = v.mean(axis=0)
        planes = []
        for f in faces:
            p = v[list(f)]
            n = np.cross(p[1] - p[0], p[2] - p[0])
            if np.linalg.norm(n) < 1e-12:
                n = np.cross(p[2] - p[0], p[3] - p[0])
            n /= np.linalg.norm(n)
            d = float(np.dot(n, p[0]))
            if np.dot(n, c) > d:
                n, d = -n, -d
            planes.append((n[0], n[1], n[2], d))
        lo, hi = v.min(axis=0), v.max(axis=0)
        return csg.convexpolyhedron(planes, (*lo, *hi))

    def _trd(self, e, name):
        x1, x2 = 0.5 * self._len(e, 'x1'), 0.5 * self._len(e, 'x2')
        y1, y2 = 0.5 * self._len(e, 'y1'), 0.5 * self._len(e, 'y2')
        hz = 0.5 * self._len(e, 'z')
        verts = [(-x1, -y1, -hz), (x1, -y1, -hz), (-x1, y1, -hz), (x1, y1, -hz),
                 (-x2, -y2, hz), (x2, -y2, hz), (-x2, y2, hz), (x2, y2, hz)]
        self.notes.append(f'trd {name}: convex polyhedron (not converted by U4Solid)')
        return self._hull(name, verts)

    def _trap(self, e, name):
        dz = 0.5 * self._len(e, 'z')
        th, ph = self._ang(e, 'theta'), self._ang(e, 'phi')
        dy1, dx1, dx2 = 0.5 * self._len(e, 'y1'), 0.5 * self._len(e, 'x1'), 0.5 * self._len(e, 'x2')
        dy2, dx3, dx4 = 0.5 * self._len(e, 'y2'), 0.5 * self._len(e, 'x3'), 0.5 * self._len(e, 'x4')
        ta1, ta2 = math.tan(self._ang(e, 'alpha1')), math.tan(self._ang(e, 'alpha2'))
        tx, ty = math.tan(th) * math.cos(ph), math.tan(th) * math.sin(ph)
        verts = []
        for z, dy, dxm, dxp, ta in ((-dz, dy1, dx1, dx2, ta1), (dz, dy2, dx3, dx4, ta2)):
            cx, cy = z * tx, z * ty
            verts += [(cx - dy * ta - dxm, cy - dy, z), (cx - dy * ta + dxm, cy - dy, z),
                      (cx + dy * ta - dxp, cy + dy, z), (cx + dy * ta + dxp, cy + dy, z)]
        self.notes.append(f'trap {name}: convex polyhedron (not converted by U4Solid)')
        return self._hull(name, verts)


# ------------------------------------------------------------------ physics tables
def _domain_energy_mev():
    return HC_EVNM / csg.domain() * 1e-6


def _sample(mat, energies):
    """G4PhysicsVector::Value: linear in energy, clamped to the end values."""
    e, v = mat[:, 0], mat[:, 1]
    order = np.argsort(e, kind='stable')
    return np.interp(energies, e[order], v[order])


def _group_velocity(rindex):
    """GROUPVEL from RINDEX pairs: c/(n + dn/dlnE) at the first energy and at the
    midpoints of the following intervals, c/n where that is not physical
    (Geant4 G4MaterialPropertiesTable::CalculateGROUPVEL)."""
    e, n = rindex[:, 0], rindex[:, 1]
    if len(e) == 1:
        return np.array([[e[0], C_LIGHT / n[0]]])
    out = []
    e0, n0, e1, n1 = e[0], n[0], e[1], n[1]
    vg = C_LIGHT / (n0 + (n1 - n0) / math.log(e1 / e0))
    if vg < 0 or vg > C_LIGHT / n0:
        vg = C_LIGHT / n0
    out.append((e0, vg))
    for i in range(2, len(e)):
        vg = C_LIGHT / (0.5 * (n0 + n1) + (n1 - n0) / math.log(e1 / e0))
        if vg < 0 or vg > C_LIGHT / (0.5 * (n0 + n1)):
            vg = C_LIGHT / (0.5 * (n0 + n1))
        out.append((0.5 * (e0 + e1), vg))
        e0, n0, e1, n1 = e1, n1, e[i], n[i]
    vg = C_LIGHT / (n1 + (n1 - n0) / math.log(e1 / e0))
    if vg < 0 or vg > C_LIGHT / n1:
        vg = C_LIGHT / n1
    out.append((e1, vg))
    return np.asarray(out)


_FINISH = {'polished': 0, 'polishedfrontpainted': 1, 'polishedbackpainted': 2, 'ground': 3,
           'groundfrontpainted': 4, 'groundbackpainted': 5}


def _finish_code(v):
    if v is None:
        return 0
    v = str(v).strip()
    if v in _FINISH:
        return _FINISH[v]
    try:
        return int(float(v))
    except ValueError:
        raise GdmlError(f'unknown surface finish {v!r}')


# ------------------------------------------------------------------ conversion
def _digest(reader, nodes, i):
    nd = nodes[i]
    parts = [nd.lv]
    for c in nd.children:
        rel = nodes[c].m2w @ np.linalg.inv(nd.m2w)
        parts.append(nodes[c].digest + ':' + ','.join(f'{x:.6g}' for x in rel.ravel()))
    return '|'.join(parts)


def load_gdml(path, freq_cut=500, max_bounce=None) -> GdmlGeometry:
    """Read a GDML file into CSGFoundry geometry arrays and SSim physics tables."""
    r = GdmlReader(path)
    if r.world not in r.volumes:
        raise GdmlError(f'world volume {r.world} not defined')

    # placed-volume tree in preorder (stree nds order)
    nodes: List[_PlacedNode] = []

    def visit(pv, lv, m2w, parent, depth, copyno):
        if depth > 256:
            raise GdmlError('volume tree deeper than 256 (recursive placement?)')
        i = len(nodes)
        nodes.append(_PlacedNode(i, pv, lv, m2w, parent, depth, copyno))
        if parent >= 0:
            nodes[parent].children.append(i)
        vol = r.volumes.get(lv)
        if vol is None:
            raise GdmlError(f'unknown volume {lv}')
        for pvname, child, local, cn in vol.children:
            visit(pvname, child, local @ m2w, i, depth + 1, cn)
    visit(r.world, r.world, np.eye(4), -1, 0, 0)
    for i in reversed(range(len(nodes))):
        nodes[i].digest = _digest(r, nodes, i)

    # materials and logical volumes (mesh index, lvid) in postorder first use
    # (U4Tree::initMaterials_r, U4Tree::initSolids_r); RINDEX decides implicit surfaces
    mat_order: List[str] = []
    lv_index: Dict[str, int] = {}

    def post(i):
        for c in nodes[i].children:
            post(c)
        m = r.volumes[nodes[i].lv].material
        if m not in mat_order:
            mat_order.append(m)
        if nodes[i].lv not in lv_index:
            lv_index[nodes[i].lv] = len(lv_index)
    post(0)
    for m in mat_order:
        if m not in r.materials:
            raise GdmlError(f'unknown material {m}')
    has_rindex = {m: 'RINDEX' in r.materials[m] for m in mat_order}

    # surfaces: border, then skin, then implicits as found
    border = {(pv1, pv2): (name, osn) for name, pv1, pv2, osn in r.border}
    skin = {lv: (name, osn) for name, lv, osn in r.skin}
    sur_names = [b[0] for b in r.border] + [s[0] for s in r.skin]
    sur_optical = {b[0]: b[3] for b in r.border}
    sur_optical.update({s[0]: s[2] for s in r.skin})
    implicit: List[str] = []
    lv_of_pv = {nd.pv: nd.lv for nd in nodes}

    def find(pre: _PlacedNode, post_: _PlacedNode):
        """U4Surface::Find(thePrePV, thePostPV)."""
        b = border.get((pre.pv, post_.pv))
        if b:
            return b[0]
        entered_daughter = post_.parent == pre.index
        first, second = (post_.lv, pre.lv) if entered_daughter else (pre.lv, post_.lv)
        s = skin.get(first) or skin.get(second)
        return s[0] if s else None

    bnd_specs: List[tuple] = []
    for nd in nodes:
        imat = r.volumes[nd.lv].material
        parent = nodes[nd.parent] if nd.parent >= 0 else None
        omat = r.volumes[parent.lv].material if parent else imat
        osur = find(parent, nd) if parent else None
        isur = find(nd, parent) if (parent and has_rindex[imat]) else None
        if osur is None and parent and has_rindex[omat] and not has_rindex[imat]:
            osur = 'Implicit_RINDEX_NoRINDEX_' + strip_name(parent.pv) + '_' + strip_name(nd.pv)
            if osur not in implicit:
                implicit.append(osur)
        if isur is None and parent and has_rindex[imat] and not has_rindex[omat]:
            isur = 'Implicit_RINDEX_NoRINDEX_' + strip_name(nd.pv) + '_' + strip_name(parent.pv)
            if isur not in implicit:
                implicit.append(isur)
        spec = (omat, osur, isur, imat)
        if spec not in bnd_specs:
            bnd_specs.append(spec)
        nd.boundary = bnd_specs.index(spec)

    # physics tables
    f = csg.Foundry()
    energies = _domain_energy_mev()
    defaults = dict(RINDEX=1.0, ABSLENGTH=1e12, RAYLEIGH=1e12, REEMISSIONPROB=0.0)
    for m in mat_order:
        props = r.materials[m]
        vals = {k: (_sample(props[k], energies) if k in props and props[k].shape[1] == 2 else d)
                for k, d in defaults.items()}
        if 'GROUPVEL' in props and props['GROUPVEL'].shape[1] == 2:
            vg = _sample(props['GROUPVEL'], energies)
        elif 'RINDEX' in props and props['RINDEX'].shape[1] == 2:
            vg = _sample(_group_velocity(props['RINDEX']), energies)
        else:
            vg = C_LIGHT
        f.add_material(m, vals['RINDEX'], vals['ABSLENGTH'], vals['RAYLEIGH'], vals['REEMISSIONPROB'],
                       group_velocity=vg)
    for name in sur_names:
        osn = sur_optical[name]
        if osn not in r.optical_surfaces:
            raise GdmlError(f'surface {name}: unknown optical surface {osn}')
        os_ = r.optical_surfaces[osn]
        props = os_['props']
        effi = _sample(props['EFFICIENCY'], energies) if 'EFFICIENCY' in props else np.zeros(len(energies))
        refl = _sample(props['REFLECTIVITY'], energies) if 'REFLECTIVITY' in props else np.zeros(len(energies))
        finish = _finish_code(os_['finish'])
        if effi.max() > 0.0:
            payload = dict(detect=effi, absorb=1.0 - effi)
        elif finish in (0, 1, 2):
            payload = dict(absorb=1.0 - refl, specular=refl)
        else:
            payload = dict(absorb=1.0 - refl, diffuse=refl)
        f.add_surface(name, osn0=(osn[:1] or '\0'), finish=finish,
                      model_value_percent=int(100.0 * os_['value']), **payload)
    for name in implicit:
        f.add_implicit_absorber(name)
    for spec in bnd_specs:
        f.boundary(*spec)

    # factorisation: subtrees repeated at least freq_cut times, outermost first
    counts: Dict[str, int] = {}
    for nd in nodes:
        counts[nd.digest] = counts.get(nd.digest, 0) + 1
    factor_digests: List[str] = []
    instance_roots: Dict[str, List[int]] = {}
    in_factor = np.zeros(len(nodes), dtype=bool)

    def mark(i):
        in_factor[i] = True
        for c in nodes[i].children:
            mark(c)
    for nd in nodes:
        if nd.index == 0 or in_factor[nd.index] or counts[nd.digest] < freq_cut:
            continue
        if nd.digest not in instance_roots:
            factor_digests.append(nd.digest)
            instance_roots[nd.digest] = []
        instance_roots[nd.digest].append(nd.index)
        mark(nd.index)

    prim_volumes: List[str] = []
    prim_boundary: List[int] = []
    solid_cache: Dict[str, csg.Node] = {}

    def tree(lv):
        s = r.volumes[lv].solid
        if s not in solid_cache:
            solid_cache[s] = r.solid(s)
        return solid_cache[s]

    def add(nd, frame):
        t = tree(nd.lv)
        m = nd.m2w @ np.linalg.inv(frame)
        if not np.allclose(m, np.eye(4), atol=1e-12):
            t = t.moved(m)
        f.add_prim(t, nd.boundary, mesh_idx=lv_index[nd.lv])
        prim_volumes.append(nd.pv)
        prim_boundary.append(nd.boundary)

    f.begin_solid('r0')
    for nd in nodes:
        if not in_factor[nd.index]:
            add(nd, np.eye(4))
    f.end_solid()
    f.add_instance(np.eye(4), 0)
    for k, dg in enumerate(factor_digests):
        roots = instance_roots[dg]
        first = nodes[roots[0]]
        f.begin_solid(f'r{k + 1}')
        sub = []

        def collect(i):
            sub.append(i)
            for c in nodes[i].children:
                collect(c)
        collect(first.index)
        for i in sub:
            add(nodes[i], first.m2w)
        f.end_solid()
        for ri in roots:
            f.add_instance(nodes[ri].m2w, k + 1)

    # scintillation ICDF from the first material with a spectrum
    phys = f.physics()
    for m in mat_order:
        props = r.materials[m]
        spec = props.get('SCINTILLATIONCOMPONENT1', props.get('FASTCOMPONENT'))
        if spec is not None and spec.shape[1] == 2 and len(spec) >= 2:
            e_mev, val = spec[:, 0], spec[:, 1]
            order = np.argsort(e_mev)

            def intensity(wl, e=e_mev[order] * 1e6, v=val[order]):
                return np.interp(HC_EVNM / wl, e, v)
            phys['icdf'] = csg.scint_icdf(intensity, hd_factor=20)
            phys['hd_factor'] = 20
            r.notes.append(f'scintillation ICDF from material {strip_name(m)}')
            break

    fmt = lambda s: strip_name(s) if s is not None else ''
    return GdmlGeometry(
        geometry=f.arrays(), physics=phys, materials=[strip_name(m) for m in mat_order],
        surfaces=[strip_name(x) for x in sur_names + implicit],
        boundaries=['/'.join(fmt(x) for x in spec) for spec in bnd_specs],
        prim_volumes=[strip_name(x) for x in prim_volumes], prim_boundary=prim_boundary,
        num_factor=len(factor_digests), notes=r.notes)
