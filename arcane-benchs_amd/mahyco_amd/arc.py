"""MaHyCo case files (.arc) for the drop-in library.

A case file configures the reference through Arcane: the Cartesian mesh
generator, the `<mahyco>` module options (Mahyco.axl:170-340), the
environments and their equation of state, the initialisation service
(`cas-model`, SOD.axl:136) and the remap service options (RemapADI.axl:
136-178).  `read_case` turns one into the mesh description and `mhg_options`
that `mhg_create` takes, and `Case.build` creates and initialises a context
on either library (the HIP one or the oracle).

Options the library does not run (the host-only branches of the reference,
other initialisation services, non-uniform mesh generators) raise
`UnsupportedOption` naming the option; elements the axl files do not declare
are listed in `Case.ignored`.  The Arcane-only settings (post-processing,
checkpoints, `nsd` sub-domain counts, ghost layers, `cartesian-sort-faces`,
`geometry`) do not change the computation on one GPU and are skipped.

    python -m mahyco_amd.arc case.arc [--n NX NY [NZ]] [--steps K] [--fp32]
"""
import argparse
import sys
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field

import numpy as np

from . import abi

SURFACES = ("XMIN", "XMAX", "YMIN", "YMAX", "ZMIN", "ZMAX")
BC_TYPES = {"Vx": abi.BC_VX, "Vy": abi.BC_VY, "Vz": abi.BC_VZ}
EOS = {"PerfectGas": abi.EOS_PERFECT_GAS, "StiffenedGas": abi.EOS_STIFFENED_GAS}
LC = {"racine-cubique-volume": abi.LC_RACINE_CUBIQUE_VOLUME, "faces-opposees": abi.LC_FACES_OPPOSEES}
SOD_CASES = (3, 4, 5, 13, 14, 15)  # SOD along x, y, z; one or two environments

# <mahyco> options (Mahyco.axl) -> mhg_options field, kind
MAHYCO_OPTIONS = {
    "schema-csts": ("schema_csts", "bool"),
    "pseudo-centree": ("pseudo_centree", "bool"),
    "deltat-min": ("deltat_min", "real"),
    "deltat-max": ("deltat_max", "real"),
    "deltat-init": ("deltat_init", "real"),
    "final-time": ("final_time", "real"),
    "cfl": ("cfl", "real"),
    "threshold": ("threshold", "real"),
    "sans-lagrange": ("sans_lagrange", "bool"),
    "with-newton": ("with_newton", "bool"),
    "with-projection": ("with_projection", "bool"),
}
REQUIRED = ("deltat-min", "deltat-max", "deltat-init", "final-time")  # no default in Mahyco.axl
# remap service options (RemapADI.axl) -> mhg_options field, kind
REMAP_OPTIONS = {
    "threshold": ("remap_threshold", "real"),
    "is-euler-scheme": ("is_euler_scheme", "bool"),
    "ordre-projection": ("ordre_projection", "int"),
    "conservation-energie-totale": ("conservation_energie_totale", "bool"),
    "projection-limiteur-id": ("projection_limiteur_id", "int"),
    "projection-pente-borne": ("projection_pente_borne", "bool"),
    "projection-pente-borne-mixte": ("projection_pente_borne_mixte", "bool"),
    "projection-limiteur-pure-id": ("projection_limiteur_pure_id", "int"),
    "projection-pente-borne-debar-fix": ("projection_pente_borne_debar_fix", "int"),
}
# remap options the library runs only at their default value
REMAP_DEFAULT_ONLY = {"calcul_flux_masse": ("int", 0)}
SKIPPED = {"material", "cartesian-sort-faces", "geometry"}


class UnsupportedOption(ValueError):
    pass


def _value(text, kind, where):
    t = (text or "").strip()
    try:
        if kind == "bool":
            if t.lower() in ("true", "1"):
                return 1
            if t.lower() in ("false", "0"):
                return 0
            raise ValueError(t)
        if kind == "int":
            return int(t)
        return float(t)
    except ValueError:
        raise ValueError(f"{where}: cannot read {t!r} as {kind}") from None


@dataclass
class Case:
    dim: int
    n: list
    origin: list
    length: list
    options: dict = field(default_factory=dict)
    envs: list = field(default_factory=list)      # (name, eos, adiabatic-cst, limit-tension)
    bcs: list = field(default_factory=list)       # (surface, type, value)
    cas_model: str = "SOD"
    cas_test: int = 0
    ignored: list = field(default_factory=list)

    def mesh(self, n=None):
        nn = list(n) if n is not None else list(self.n)
        if len(nn) != self.dim:
            raise ValueError(f"mesh override {nn} does not have {self.dim} entries")
        return abi.mesh_desc(self.dim, nn, self.origin, self.length)

    def mhg_options(self, lib, **override):
        opt = lib.options()
        for k, v in {**self.options, **override}.items():
            setattr(opt, k, v)
        abi.set_envs(opt, [(eos, g, t) for _, eos, g, t in self.envs])
        abi.set_bcs(opt, self.bcs)
        return opt

    def build(self, lib, n=None, **override):
        """A context for this case on `lib`, initialised by the case's
        initialisation service.  n: cells per direction instead of the
        file's; override: mhg_options fields (precision, ...)."""
        nn = list(n) if n is not None else list(self.n)
        s = abi.Simulation(lib, self.mesh(nn), self.mhg_options(lib, **override))
        if self.dim == 2:
            s.init_sod(self.cas_test)
        else:  # the library initialises 3D states cell by cell
            s.init_cells(*self.sod_cells(nn))
        return s

    def sod_cells(self, n):
        """Environment, density and pressure of each cell of the SOD case
        (SODService.cc:11-39): left state (1, 1) where the cell centre along
        the case's axis is below 0.5, right state (0.125, 0.1) beyond; cases
        13-15 put the right state in the second environment."""
        axis = self.cas_test - (3 if self.cas_test < 13 else 13)
        idx = np.indices(tuple(reversed(n))).reshape(len(n), -1)[::-1]  # i fastest
        r = self.origin[axis] + (idx[axis] + 0.5) * (self.length[axis] / n[axis])
        left = r < 0.5
        env = np.where(left | (self.cas_test < 13), 0, 1).astype(np.int32)
        return env, np.where(left, 1.0, 0.125), np.where(left, 1.0, 0.1)


def _child_text(el, tag):
    c = el.find(tag)
    return None if c is None else c.text


def _mesh(root):
    gen = root.find("mesh/meshgenerator")
    if gen is None:
        raise UnsupportedOption("mesh: only <meshgenerator> meshes are read")
    cart = gen.find("cartesian")
    if cart is None:
        names = [c.tag for c in gen]
        raise UnsupportedOption(f"mesh generator {names}: only <cartesian> is read")
    origin = [float(v) for v in (_child_text(cart, "origine") or "0 0 0").split()]
    n, length = [], []
    for axis, tag in enumerate(("lx", "ly", "lz")):
        el = cart.find(tag)
        if el is None:
            break
        n.append(int(el.get("n" + tag[1])))
        length.append(float(el.text))
        pr = float(el.get("pr" + tag[1], "1.0"))
        if pr != 1.0:
            raise UnsupportedOption(f"mesh {tag} pr{tag[1]}={pr}: only uniform spacing (1.0) is run")
    dim = len(n)
    if dim not in (2, 3):
        raise UnsupportedOption(f"mesh of dimension {dim}")
    return dim, n, (origin + [0.0, 0.0, 0.0])[:dim], length


def _apply(table, el, options, where):
    name = el.tag
    fld, kind = table[name]
    options[fld] = _value(el.text, kind, f"{where}/{name}")


def read_case(path_or_text):
    """Parse a MaHyCo .arc case (a path, or the XML text itself)."""
    text = path_or_text
    if "<" not in path_or_text:
        with open(path_or_text, "r", encoding="utf-8") as f:
            text = f.read()
    root = ET.fromstring(text)
    dim, n, origin, length = _mesh(root)
    case = Case(dim=dim, n=n, origin=origin, length=length)
    mah = root.find("mahyco")
    if mah is None:
        raise ValueError("no <mahyco> element")
    seen = set()
    for el in mah:
        tag = el.tag
        if not isinstance(tag, str):  # comments
            continue
        if tag in MAHYCO_OPTIONS:
            _apply(MAHYCO_OPTIONS, el, case.options, "mahyco")
            seen.add(tag)
        elif tag == "longueur-caracteristique":
            v = (el.text or "").strip()
            if v not in LC:
                raise ValueError(f"mahyco/longueur-caracteristique: {v!r}")
            case.options["longueur_caracteristique"] = LC[v]
        elif tag == "environment":
            name = _child_text(el, "name") or f"env{len(case.envs)}"
            mats = el.findall("material")
            if len(mats) > 1:
                raise UnsupportedOption(f"environment {name}: {len(mats)} materials (one per environment is run)")
            eos = el.find("eos-model")
            kind = eos.get("name") if eos is not None else None
            if kind not in EOS:
                raise UnsupportedOption(f"environment {name}: eos-model {kind!r}")
            g = _value(_child_text(eos, "adiabatic-cst") or "1.4", "real", f"{name}/adiabatic-cst")
            t = _value(_child_text(eos, "limit-tension") or "0.0", "real", f"{name}/limit-tension")
            case.envs.append((name, EOS[kind], g, t))
        elif tag == "cas-model":
            case.cas_model = el.get("name", "SOD")
            ct = _child_text(el, "cas-test")
            if ct is not None:
                case.cas_test = int(ct)
        elif tag == "cas-test":  # older layout: directly under <mahyco>
            case.cas_test = int(el.text)
        elif tag == "remap":
            if el.get("name", "RemapADI") != "RemapADI":
                raise UnsupportedOption(f"remap service {el.get('name')!r}: RemapADI is run")
            for r in el:
                if isinstance(r.tag, str):
                    _remap_option(case, r)
        elif tag in REMAP_OPTIONS or tag in REMAP_DEFAULT_ONLY:  # older layout
            _remap_option(case, el)
        elif tag == "boundary-condition":
            surf = (_child_text(el, "surface") or "").strip()
            typ = (_child_text(el, "type") or "").strip()
            if surf not in SURFACES or typ not in BC_TYPES:
                raise ValueError(f"boundary-condition {surf!r} {typ!r}")
            val = _value(_child_text(el, "value") or "0", "real", f"boundary-condition {surf}")
            case.bcs.append((surf, BC_TYPES[typ], val))
        elif tag in SKIPPED:
            pass
        else:
            case.ignored.append(tag)
    missing = [r for r in REQUIRED if r not in seen]
    if missing:
        raise ValueError(f"mahyco: missing required options {missing}")
    if case.cas_model != "SOD" or case.cas_test not in SOD_CASES:
        raise UnsupportedOption(f"cas-model {case.cas_model} cas-test {case.cas_test}: SOD cases {SOD_CASES} are run")
    if not case.envs:
        raise ValueError("no <environment>")
    if dim == 2:  # a 2D mesh has no z faces
        case.bcs = [b for b in case.bcs if b[0] not in ("ZMIN", "ZMAX")]
    return case


def _remap_option(case, el):
    tag = el.tag
    if tag in REMAP_OPTIONS:
        _apply(REMAP_OPTIONS, el, case.options, "remap")
    elif tag in REMAP_DEFAULT_ONLY:
        kind, default = REMAP_DEFAULT_ONLY[tag]
        if _value(el.text, kind, f"remap/{tag}") != default:
            raise UnsupportedOption(f"remap/{tag}={el.text.strip()}: only the default {default} is run")
    else:
        case.ignored.append("remap/" + tag)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("case")
    ap.add_argument("--fp32", action="store_true", help="FP32-bulk precision form")
    ap.add_argument("--n", type=int, nargs="+", help="cells per direction instead of the file's")
    ap.add_argument("--steps", type=int, default=1 << 30)
    a = ap.parse_args(argv)
    case = read_case(a.case)
    from . import load_gpu
    s = case.build(load_gpu(), n=a.n, precision=abi.PREC_FP32_BULK if a.fp32 else abi.PREC_FP64)
    n = s.run(a.steps)
    st = s.download()
    print(f"{a.case}: {case.dim}D {a.n or case.n}, {len(case.envs)} environments, cas-test {case.cas_test}: "
          f"{n} iterations, t = {st['global_time']:.6g}, dt = {st['global_deltat']:.3g}")
    if case.ignored:
        print("ignored elements:", ", ".join(case.ignored))
    s.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
