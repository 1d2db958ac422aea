"""CVFES ``.cfg`` reader: a ConfigObj-compatible subset parser and the ``CVConfig`` tree.

CVFES reads its nested INI files with ConfigObj (``cvfes.py:38``, ``cvconfig.py:166-219``), which
is not installed here.  :class:`Section` implements the part of ConfigObj the configuration classes
use: nested ``[a]``/``[[b]]``/``[[[c]]]`` sections, ``key = value`` with quoted strings,
comma-separated lists, ``#`` comments, ``.sections``, ``.as_int/.as_float/.as_bool``, ``.parent``
and ``.name``.

The classes below mirror ``cvconfig.py:16-219`` attribute for attribute.  Keys the reference
requires and a solid-only run does not use get defaults instead of raising.  New optional keys
(all additive):

* solver: ``beta`` (Rayleigh stiffness damping), ``stiffness_method`` (``stored`` | ``matfree``),
  ``mass_lumping`` (``cpu`` | ``gpu``), ``pressure_waveform`` (``ramp`` | ``sine``) with
  ``pressure_mean``, ``pressure_amplitude`` (relative) and ``pressure_period``;
* wall mesh: ``generator = tube`` with ``ntheta``, ``nz``, ``radius``, ``radius_out``, ``length``,
  ``stenosis`` (synthetic tube instead of a VTK file);
* solid equation: ``Youngs Modulus model = gmrf_element`` with ``gmrf seed`` and ``gmrf cov``.
"""
from __future__ import annotations

import re

__all__ = ["Section", "parse_cfg", "read_cfg", "CVConfig", "SolverConfig", "EquationConfig",
           "MeshConfig", "FaceConfig"]

_TRUE = {"true", "yes", "on", "1"}
_FALSE = {"false", "no", "off", "0"}


class Section(dict):
    """A ConfigObj-like section (ordered dict with typed accessors)."""

    def __init__(self, name="", parent=None):
        super().__init__()
        self.name = name
        self.parent = parent
        self.sections = []
        self.scalars = []

    def as_int(self, key):
        return int(self[key])

    def as_float(self, key):
        return float(self[key])

    def as_bool(self, key):
        v = self[key]
        if isinstance(v, bool):
            return v
        s = str(v).strip().lower()
        if s in _TRUE:
            return True
        if s in _FALSE:
            return False
        raise ValueError(f"Value {v!r} of {key!r} is neither True nor False")

    def as_list(self, key):
        v = self[key]
        return v if isinstance(v, list) else [v]


def _strip_comment(line):
    out, quote = [], None
    for ch in line:
        if quote:
            out.append(ch)
            if ch == quote:
                quote = None
        elif ch in ("'", '"'):
            quote = ch
            out.append(ch)
        elif ch == "#":
            break
        else:
            out.append(ch)
    return "".join(out).strip()


def _parse_value(text):
    """ConfigObj value: unquoted commas make a list, quotes are stripped."""
    items, cur, quote, has_comma = [], [], None, False
    for ch in text:
        if quote:
            if ch == quote:
                quote = None
            else:
                cur.append(ch)
        elif ch in ("'", '"'):
            quote = ch
        elif ch == ",":
            items.append("".join(cur).strip())
            cur = []
            has_comma = True
        else:
            cur.append(ch)
    last = "".join(cur).strip()
    if not has_comma:
        return last
    if last != "":
        items.append(last)
    return items


_SEC = re.compile(r"^(\[+)\s*(.*?)\s*(\]+)$")


def parse_cfg(text: str) -> Section:
    """Parse ConfigObj-style text into a :class:`Section` tree."""
    root = Section("", None)
    stack = [root]
    for lineno, raw in enumerate(text.splitlines(), 1):
        line = _strip_comment(raw)
        if not line:
            continue
        m = _SEC.match(line)
        if m:
            depth = len(m.group(1))
            if depth != len(m.group(3)):
                raise ValueError(f"line {lineno}: unbalanced section brackets")
            if depth > len(stack):
                raise ValueError(f"line {lineno}: section nested too deep")
            parent = stack[depth - 1]
            sec = Section(m.group(2), parent)
            parent[m.group(2)] = sec
            parent.sections.append(m.group(2))
            del stack[depth:]
            stack.append(sec)
            continue
        if "=" not in line:
            raise ValueError(f"line {lineno}: expected 'key = value'")
        key, value = line.split("=", 1)
        key = key.strip().strip("'\"")
        sec = stack[-1]
        sec[key] = _parse_value(value.strip())
        sec.scalars.append(key)
    return root


def read_cfg(path) -> Section:
    with open(path, "r") as fh:
        return parse_cfg(fh.read())


def _get(sec, key, conv, default):
    if key not in sec:
        return default
    return conv(sec, key)


_f = Section.as_float
_i = Section.as_int
_b = Section.as_bool


class Config:
    def __init__(self, config):
        self.name = config.get("name")
        self.file_path = config.get("file_path")


class FaceConfig(Config):
    pass


class MeshConfig(Config):
    """``cvconfig.py:32-46`` plus the synthetic-tube keys."""

    def __init__(self, meshSection):
        Config.__init__(self, meshSection)
        self.domainId = _get(meshSection, "domain_id", _i, 0)
        self.faces = []
        if "faces" in meshSection:
            facesConfig = meshSection["faces"]
            for section in facesConfig.sections:
                if section.startswith("face"):
                    self.faces.append(FaceConfig(facesConfig[section]))
        self.generator = meshSection.get("generator")
        self.ntheta = _get(meshSection, "ntheta", _i, None)
        self.nz = _get(meshSection, "nz", _i, None)
        self.radius = _get(meshSection, "radius", _f, None)
        self.radius_out = _get(meshSection, "radius_out", _f, None)
        self.length = _get(meshSection, "length", _f, None)
        self.stenosis = _get(meshSection, "stenosis", _f, 0.0)


class ConditionConfig:
    def getProp(self, config, key, required=True):
        if not required and key not in config:
            return None
        sub = config[key]
        if "uniform_value" in sub:
            return sub.as_float("uniform_value")
        if "uniform_func" in sub:
            return sub["uniform_func"]
        if "file_name" in sub:
            return sub["file_name"]
        raise KeyError(f"no value for {key}")


class InitialConditionsConfig(ConditionConfig):
    def __init__(self, sec, name):
        if name == "fluid":
            self.acceleration = self.getProp(sec, "acceleration", False)
            self.velocity = self.getProp(sec, "velocity", False)
            self.pressure = self.getProp(sec, "pressure", False)
        else:
            self.velocity = self.getProp(sec, "velocity")
            self.displacement = self.getProp(sec, "displacement")


class BoundaryConditionsConfig(ConditionConfig):
    def __init__(self, sec, name):
        if name == "fluid":
            self.inletVelocity = self.getProp(sec, "inlet_velocity", False)
            self.outletH = self.getProp(sec, "outlet_h", False)
            self.parabolicInlet = _get(sec, "parabolic_inlet", _b, False)
        else:
            disp = self.getProp(sec, "displacement", False)
            self.bdyDisplacement = 0.0 if disp is None else disp


class EquationConfig:
    """``cvconfig.py:98-136`` plus the GMRF modulus keys."""

    def __init__(self, sec):
        self.name = sec["name"]
        if self.name == "fluid":
            self.dviscosity = _get(sec, "dynamic_viscosity", _f, None)
            self.density = _get(sec, "density", _f, None)
            self.f = _get(sec, "external_force", _f, 0.0)
        else:
            self.thickness = sec.as_float("thickness")
            self.density = sec.as_float("density")
            self.E = sec.as_float("Youngs Modulus")
            self.v = sec.as_float("Poissons Ratio")
            self.damp = _get(sec, "Damp", _f, 0.0)
            self.thicknessFilename = sec.get("thickness Filename")
            self.sigmaThickness = _get(sec, "thickness sigma", _f, 0.0)
            self.rhoThickness = _get(sec, "thickness rho", _f, 3.7)
            self.YoungsModulusFilename = sec.get("Youngs Modulus Filename")
            self.sigmaE = _get(sec, "Youngs Modulus sigma", _f, 0.0)
            self.rhoE = _get(sec, "Youngs Modulus rho", _f, 3.7)
            self.EModel = sec.get("Youngs Modulus model")
            self.gmrfSeed = _get(sec, "gmrf seed", _i, 0)
            self.gmrfCov = _get(sec, "gmrf cov", _f, 0.10)
        self.initialConditions = (InitialConditionsConfig(sec["initial conditions"], self.name)
                                  if "initial conditions" in sec else None)
        self.boundaryConditions = (BoundaryConditionsConfig(sec["boundary conditions"], self.name)
                                   if "boundary conditions" in sec else None)


class SolverConfig:
    """``cvconfig.py:139-163`` plus the solid-step keys."""

    def __init__(self, sec):
        self.method = sec.get("method", "transient GPU")
        self.time = _get(sec, "time", _f, 0.0)
        self.dt = sec.as_float("dt")
        self.endtime = sec.as_float("endtime")
        self.tolerance = _get(sec, "tolerance", _f, None)
        self.rho_infinity = _get(sec, "rho_infinity", _f, None)
        self.imax = _get(sec, "imax", _i, None)
        self.ci = _get(sec, "ci", _f, None)
        self.dt_f = _get(sec, "dt_f", _f, None)
        self.update_interval = _get(sec, "stiffness_update_interval", _i, 1 << 62)
        self.kernel_autotune = _get(sec, "kernel_autotune", _b, True)
        self.finite_check_every = _get(sec, "finite_check_every", _i, 100)
        self.constant_pressure = _get(sec, "constant_pressure", _f, 0.0)
        self.constant_T = _get(sec, "constant_T", _f, 0.0)
        self.restart = _get(sec, "restart", _b, False)
        self.restartFilename = sec.get("restart_displacement_file")
        self.restartTimestep = _get(sec, "restart_timestep", _i, 0)
        self.saveStressFilename = None
        self.saveResNum = None
        # additive keys of the MI355X solid step
        self.beta = _get(sec, "beta", _f, 0.0)
        self.stiffness_method = sec.get("stiffness_method", "stored")
        self.mass_lumping = sec.get("mass_lumping", "cpu")
        self.pressure_waveform = sec.get("pressure_waveform", "ramp")
        self.pressure_mean = _get(sec, "pressure_mean", _f, self.constant_pressure)
        self.pressure_amplitude = _get(sec, "pressure_amplitude", _f, 0.0)
        self.pressure_period = _get(sec, "pressure_period", _f, 1.0)


class CVConfig:
    """Top-level configuration (``cvconfig.py:166-219``)."""

    def __init__(self, config: Section):
        self.ndim = _get(config, "dimensions", _i, 3)
        self.nSmp = _get(config, "sample_num", _i, 1)
        self.regenerate_samples = _get(config, "regenerate_samples", _b, False)
        self.name = config.get("name", "spCy")
        self.meshes = {}
        if "meshes" in config:
            mc = config["meshes"]
            for section in mc.sections:
                if section.startswith("mesh"):
                    self.meshes[mc[section]["name"]] = MeshConfig(mc[section])
        self.equations = {}
        if "equations" in config:
            ec = config["equations"]
            for section in ec.sections:
                if section.startswith("equation"):
                    self.equations[ec[section]["name"]] = EquationConfig(ec[section])
        self.solver = SolverConfig(config["solver"])
        self.solver.saveStressFilename = config.get("save_stress_filename")
        self.solver.saveResNum = _get(config, "save_result_num", _i, 0)
        self.solver.exportBdyStressFilename = config.get("export_stress_filename")
        self.solver.useConstantStress = _get(config, "use_constant_stress", _b, False)
        if "wall" in self.meshes:
            self.meshes["wall"].nSmp = self.nSmp
            self.meshes["wall"].regenerate_samples = self.regenerate_samples
        self.solver.nSmp = self.nSmp
