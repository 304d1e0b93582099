"""Test-side reader for ARTSS XML input files.

Mirrors the subset of the reference's ``Settings::parse_settings``
(src/utility/settings/Settings.cpp:126-970) that the hot path needs, so
tests can hand the same configuration to the CPU oracle (which takes a plain
descriptor) and to the product (which parses the XML itself in C++).
Numbers go through ``float``/``int`` (correctly rounded, like ``std::stod``);
booleans are the literal "Yes" (Settings.cpp:20).
"""
import xml.etree.ElementTree as ET

FIELD_TYPES = {"rho": 0, "u": 1, "v": 2, "w": 3, "p": 4, "T": 5, "nu": 6}
PATCHES = {"left": 0, "right": 1, "bottom": 2, "top": 3, "front": 4, "back": 5}
BCS = {"neumann": 0, "dirichlet": 1, "periodic": 2}
AXES = {"x": 0, "y": 1, "z": 2}


def axes(text):
    """Mapping::match_axis (Mapping.cpp:17-23): entries matched without
    regard to case; one that names no axis ("xyz") is UNKNOWN_AXIS, which no
    caller looks for, so it is dropped; each axis is kept once"""
    out = []
    for a in text.split(","):
        i = AXES.get(a.lower())
        if i is not None and i not in out:
            out.append(i)
    return out


def _values(el):
    """Settings.cpp map_parameter_section: attributes + child element texts."""
    vals = dict(el.attrib)
    for c in el:
        if c.text is not None:
            vals[c.tag] = c.text
    return vals


def _f(vals, key, default=None):
    if key in vals:
        return float(vals[key])
    if default is None:
        raise KeyError(key)
    return default


def _yes(vals, key, default=False):
    return vals[key] == "Yes" if key in vals else default


def _boundary(el):
    v = dict(el.attrib)
    bc = BCS[v["type"]]
    val = 0.0 if bc == 2 else float(v["value"])
    fields = [FIELD_TYPES[s] for s in v["field"].split(",")]
    patches = [PATCHES[s] for s in v["patch"].split(",")]
    return fields, patches, bc, val


def _diffusion(el):
    v = _values(el)
    d = {"type": v["type"], "fields": v.get("field", "")}
    if v["type"] in ("Jacobi", "ColoredGaussSeidel"):
        d.update(max_iter=int(v["max_iter"]), tol_res=float(v["tol_res"]), w=float(v["w"]))
    return d


def parse(path_or_text):
    text = path_or_text
    if not path_or_text.lstrip().startswith("<"):
        with open(path_or_text) as fh:
            text = fh.read()
    root = ET.fromstring(text)
    cfg = {}
    phys = _values(root.find("physical_parameters"))
    solver_el = root.find("solver")
    desc = solver_el.attrib["description"]
    cfg["solver"] = desc
    cfg["dt"] = _f(phys, "dt")
    cfg["t_end"] = _f(phys, "t_end")
    cfg["nu"] = _f(phys, "nu", 0.0)
    cfg["kappa"] = _f(phys, "kappa", 0.0)
    cfg["beta"] = _f(phys, "beta", 3.34e-3)
    cfg["g"] = _f(phys, "g", -9.81)
    cfg["gamma"] = _f(phys, "gamma", 0.0)

    dom = _values(root.find("domain_parameters"))
    cfg["domain"] = dict(
        X1=_f(dom, "X1"), X2=_f(dom, "X2"), Y1=_f(dom, "Y1"), Y2=_f(dom, "Y2"),
        Z1=_f(dom, "Z1"), Z2=_f(dom, "Z2"),
        nx=int(dom["nx"]), ny=int(dom["ny"]), nz=int(dom["nz"]))
    if _yes(dom, "enable_computational_domain"):
        raise ValueError("computational sub-domains are not supported")

    for name in ("advection", "diffusion", "turbulence", "source", "pressure", "temperature", "concentration"):
        el = solver_el.find(name)
        if el is None:
            continue
        if name == "diffusion":
            cfg["diffusion"] = _diffusion(el)
        elif name == "advection":
            cfg["advection"] = dict(el.attrib)
        elif name == "turbulence":
            v = _values(el)
            cfg["turbulence"] = {"type": v["type"], "cs": float(v.get("Cs", "0"))}
        elif name == "source":
            v = _values(el)
            s = {"type": v["type"], "force_fct": v["force_fct"],
                 "dir": axes(v["dir"])}
            if v["force_fct"] == "Buoyancy":
                s["use_init_values"] = _yes(v, "use_init_values")
                if not s["use_init_values"]:
                    s["ambient"] = float(v["ambient_temperature_value"])
            cfg["source"] = s
        elif name == "pressure":
            v = _values(el)
            cfg["pressure"] = dict(n_level=int(v["n_level"]), n_cycle=int(v["n_cycle"]),
                                   max_cycle=int(v["max_cycle"]), n_relax=int(v["n_relax"]),
                                   tol_res=float(v["tol_res"]), smoother=_diffusion(el.find("diffusion")))
        elif name == "temperature":
            t = {"diffusion": _diffusion(el.find("diffusion"))}
            turb = el.find("turbulence")
            t["turbulence"] = turb is not None and turb.attrib.get("include") == "Yes"
            if t["turbulence"]:
                t["prandtl"] = float(_values(turb)["Pr_T"])
            src = _values(el.find("source"))
            t["source"] = {"type": src["type"], "temp_fct": src["temp_fct"],
                           "dir": axes(src["dir"]),
                           "dissipation": _yes(src, "dissipation"), "random": _yes(src, "random")}
            if src["temp_fct"] == "Gauss":
                t["source"]["gauss"] = [float(src[k]) for k in
                                        ("HRR", "cp", "x0", "y0", "z0", "sigma_x", "sigma_y", "sigma_z")]
                t["source"]["tau"] = float(src["tau"])
            cfg["temperature"] = t
        elif name == "concentration":
            c = {"advection": dict(el.find("advection").attrib), "diffusion": _diffusion(el.find("diffusion"))}
            turb = el.find("turbulence")
            c["turbulence"] = turb is not None and turb.attrib.get("include") == "Yes"
            if c["turbulence"]:
                c["schmidt"] = float(_values(turb)["Sc_T"])
            src = _values(el.find("source"))
            c["source"] = {"type": src["type"], "con_fct": src["con_fct"], "dir": axes(src["dir"])}
            if src["con_fct"] == "Gauss":
                c["source"]["gauss"] = [float(src[k]) for k in
                                        ("HRR", "cp", "x0", "y0", "z0", "sigma_x", "sigma_y", "sigma_z")]
                c["source"]["tau"] = float(src["tau"])
            else:
                c["source"]["cube"] = [float(src[a + "_start"]) for a in "xyz"] + \
                    [float(src[a + "_end"]) for a in "xyz"] + [float(src["val"])]
            cfg["concentration"] = c

    cfg["boundaries"] = [_boundary(b) for b in root.find("boundaries") if b.tag == "boundary"]
    obs = root.find("obstacles")
    cfg["obstacles"] = []
    if obs is not None and obs.attrib.get("enabled") == "Yes":
        for o in obs:
            g = dict(o.find("geometry").attrib)
            box = [float(g[k]) for k in ("ox1", "ox2", "oy1", "oy2", "oz1", "oz2")]
            cfg["obstacles"].append({"name": o.attrib["name"], "box": box,
                                     "boundaries": [_boundary(b) for b in o if b.tag == "boundary"]})
    ic = root.find("initial_conditions")
    civ = _values(ic)
    cfg["ic"] = {"usr_fct": civ["usr_fct"], "values": civ, "random": _yes(civ, "random")}
    levels = 0
    if "NS" in desc or desc == "PressureSolver":
        levels = cfg["pressure"]["n_level"]
    cfg["levels"] = levels
    return cfg
