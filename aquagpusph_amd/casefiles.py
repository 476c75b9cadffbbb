"""Write a case (from aquagpusph_amd.cases) as reference-format XML files.

The layout is the one of the reference examples (examples/*/templates):
Main.xml includes the presets, declares the case variables and one
ParticlesSet per particle group, loading a FastASCII particle file.  Values
are written with 9 significant digits so that float32 data round-trips
exactly.
"""
import os

import numpy as np

PBI_PRESETS = ("resources/Presets/basic.xml",
               "resources/Presets/basic/time_scheme/improved_euler.xml",
               "resources/Presets/basic/domain.xml",
               "resources/Presets/cfd.xml",
               "resources/Presets/cfd/deltaSPH.xml",
               "resources/Presets/cfd/BI.xml",
               "resources/Presets/cfd/variableTimeStep.xml")

PGP_PRESETS = ("resources/Presets/basic.xml",
               "resources/Presets/basic/time_scheme/improved_euler.xml",
               "resources/Presets/basic/domain.xml",
               "resources/Presets/cfd.xml",
               "resources/Presets/cfd/deltaSPH.xml",
               "resources/Presets/cfd/GP.xml",
               "resources/Presets/cfd/variableTimeStep.xml")

# a periodic box of fluid (no walls, no domain)
PPER_PRESETS = ("resources/Presets/basic.xml",
                "resources/Presets/basic/time_scheme/improved_euler.xml",
                "resources/Presets/cfd.xml",
                "resources/Presets/cfd/deltaSPH.xml",
                "resources/Presets/cfd/periodic.xml",
                "resources/Presets/cfd/variableTimeStep.xml")

PBIE_PRESETS = ("resources/Presets/basic.xml",
                "resources/Presets/basic/time_scheme/improved_euler.xml",
                "resources/Presets/basic/domain.xml",
                "resources/Presets/cfd.xml",
                "resources/Presets/cfd/deltaSPH.xml",
                "resources/Presets/cfd/BIe.xml",
                "resources/Presets/cfd/variableTimeStep.xml")

FIELDS = ("r", "normal", "tangent", "u", "dudt", "rho", "drhodt", "m",
          "imove")


def fields_of(case):
    """Particle-file columns of a case (associations for ghost walls)."""
    return FIELDS + (("associations",) if "associations" in case else ())


def _f(x):
    return "%.9g" % float(np.float32(x))


def _vec(v, dims):
    vs = 2 if dims == 2 else 4
    out = [0.0] * vs
    for k in range(min(len(v), vs)):
        out[k] = float(v[k])
    return ", ".join(_f(x) for x in out)


def write_particles(case, path):
    dims = case["dims"]
    vs = 2 if dims == 2 else 4
    N = int(case["N"])
    fields = fields_of(case)
    cols = []
    for k in fields:
        v = case.get(k)
        if k in ("r", "normal", "tangent", "u", "dudt"):
            a = np.zeros((N, vs), np.float32) if v is None else \
                np.asarray(v, np.float32).reshape(N, vs)
            cols.append(a)
        elif k in ("imove", "associations"):
            cols.append(np.asarray(v, np.int64).reshape(N, 1))
        else:
            a = np.zeros(N, np.float32) if v is None else \
                np.asarray(v, np.float32)
            cols.append(a.reshape(N, 1))
    fmt = []
    for k, c in zip(fields, cols):
        fmt.append(" ".join(["%d" if k in ("imove", "associations")
                             else "%.9g"] * c.shape[1]))
    data = np.concatenate([c.astype(np.float64) for c in cols], axis=1)
    np.savetxt(path, data, fmt=", ".join(fmt),
               header="aquagpusph_amd particles: " + ", ".join(fields))


def write_case(case, folder, presets=PBI_PRESETS, load=True, extra_xml="",
               save=None, timing=None, portals=None):
    """Write Main.xml (+ Fluid.dat when load is True) and return its path.

    With load=False the ParticlesSet has no Load element and the arrays are
    uploaded afterwards through Simulation.set (large cases).  `save` =
    (format, file pattern, fields) adds a <Save> to the set; `timing` is a
    list of (name, type, value) <Timing> options, e.g. ("End", "Steps", 10),
    ("Output", "IPF", 5).  `portals` is a list of (prefix, out_r, in_r, n)
    plane portals: cfd/portal.xml is included once per prefix, after the
    presets, and set_scalar tools set portal_out_r, portal_in_r and portal_n
    before each of that prefix's mirror and teleport tools (the layout of
    examples/2D/taylor_green/src/templates/BCs.xml)."""
    os.makedirs(folder, exist_ok=True)
    dims = case["dims"]
    p = case["params"]
    lines = ['<?xml version="1.0" ?>', "<sphInput>"]
    for pr in presets:
        lines.append('    <Include file="%s" />' % pr)
    for px, _, _, _ in portals or ():
        lines.append('    <Include file="resources/Presets/cfd/portal.xml" '
                     'prefix="%s" />' % px)
    if portals:
        lines.append("    <Tools>")
        for px, out_r, in_r, n in portals:
            for var, v in (("portal_out_r", out_r), ("portal_in_r", in_r),
                           ("portal_n", n)):
                lines.append('        <Tool action="insert" before="%scfd '
                             'portal* mirror*" type="set_scalar" name="%s%s" '
                             'in="%s" value="%s"/>'
                             % (px, px, var, var, _vec(v, dims)))
        lines.append("    </Tools>")
    lines.append("    <Variables>")
    var = [("g", "vec", _vec(p["g"], dims)),
           ("h", "float", _f(p["h"])),
           ("cs", "float", _f(p["cs"])),
           ("courant", "float", _f(p["courant"])),
           ("dt_min", "float", _f(p.get("dt_min", 0.0))),
           ("dt_Ma", "float", _f(p.get("dt_Ma", 0.1))),
           ("p0", "float", _f(p.get("p0", 0.0)))]
    if "hfac" in p:
        var.append(("hfac", "float", _f(p["hfac"])))
    if "domain_min" in p:
        var.append(("domain_min", "vec", _vec(p["domain_min"], dims)))
        var.append(("domain_max", "vec", _vec(p["domain_max"], dims)))
    if "periodic" in p:
        # the box of Presets/cfd/periodic.xml
        var.append(("periodic_min", "vec", _vec(p["periodic"]["min"], dims)))
        var.append(("periodic_L", "vec", _vec(p["periodic"]["L"], dims)))
    for name, t, v in var:
        lines.append('        <Variable name="%s" type="%s" value="%s" />'
                     % (name, t, v))
    lines.append("    </Variables>")
    if "sets" in case:
        # one ParticlesSet per case of a batch, particles contiguous
        o = 0
        for k, (n, sc) in enumerate(case["sets"]):
            lines.append('    <ParticlesSet n="%d">' % int(n))
            for name, v in sc.items():
                txt = (_vec(v, dims) if isinstance(v, (tuple, list))
                       else _f(v))
                lines.append('        <Scalar name="%s" value="%s" />'
                             % (name, txt))
            if load:
                sub_case = {key: (val[o:o + n] if key in
                                  ("r", "normal", "tangent", "u", "dudt",
                                   "rho", "drhodt", "m", "imove") else val)
                            for key, val in case.items()}
                sub_case["N"] = n
                fname = "Fluid_%d.dat" % k
                write_particles(sub_case, os.path.join(folder, fname))
                lines.append('        <Load format="FastASCII" file="%s" '
                             'fields="%s" />'
                             % (fname, ", ".join(fields_of(case))))
            lines.append("    </ParticlesSet>")
            o += n
    else:
        lines.append('    <ParticlesSet n="%d">' % int(case["N"]))
        for k in ("refd", "visc_dyn", "delta"):
            lines.append('        <Scalar name="%s" value="%s" />'
                         % (k, _f(p[k][0])))
        if load:
            write_particles(case, os.path.join(folder, "Fluid.dat"))
            lines.append('        <Load format="FastASCII" file="Fluid.dat" '
                         'fields="%s" />' % ", ".join(fields_of(case)))
        if save is not None:
            fmt, pattern, fl = save
            lines.append('        <Save format="%s" file="%s" fields="%s" />'
                         % (fmt, pattern, ", ".join(fl)))
        lines.append("    </ParticlesSet>")
    if timing:
        lines.append("    <Timing>")
        for name, typ, val in timing:
            lines.append('        <Option name="%s" type="%s" value="%s" />'
                         % (name, typ, val))
        lines.append("    </Timing>")
    if extra_xml:
        lines.append(extra_xml)
    lines.append("</sphInput>")
    path = os.path.join(folder, "Main.xml")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path


def upload(sim, case):
    """Set the particle arrays of a case on a Simulation built with
    write_case(load=False)."""
    dims = case["dims"]
    vs = 2 if dims == 2 else 4
    N = int(case["N"])
    for k in ("r", "normal", "tangent", "u", "dudt"):
        v = case.get(k)
        if v is not None:
            sim.set(k, np.asarray(v, np.float32).reshape(N, vs))
    for k in ("rho", "drhodt", "m"):
        v = case.get(k)
        if v is not None:
            sim.set(k, np.asarray(v, np.float32))
    sim.set("imove", np.asarray(case["imove"], np.int32))
    if "associations" in case:
        sim.set("associations", np.asarray(case["associations"], np.uint32))
