"""Python handle on the runtime (aq_sim_* in include/aquagpusph_amd.h).

    sim = Simulation("Main.xml", dims=3)
    sim.step(100)              # eager: walk the tool list per step
    sim.step(100, graph=True)  # replay one captured HIP graph per step
    r = sim.get("r")           # numpy view of a variable, reference layout

The C++ runtime does all the work; this module only marshals arrays.
"""
import ctypes
import os

import numpy as np

from . import _lib

_NP = {"f": np.float32, "i": np.int32, "u": np.uint32, "d": np.float64}


def _base(type_name):
    t = type_name.rstrip("*")
    if t.startswith("uivec") or t == "unsigned int":
        return "u"
    if t.startswith("ivec") or t == "int":
        return "i"
    if t.startswith("dvec") or t == "double":
        return "d"
    return "f"


def problem_tools(xml_path, dims=3, keys=False):
    """The (name, type) tool list a case file resolves to, in execution
    order, parsed on the host (aq_xml_resolve: no GPU needed).  With
    keys=True each entry is (name, type, key): the registry key
    "<script under resources/Scripts>:<entry_point>" of a kernel tool, the
    `in` variable of the others."""
    L = _lib.lib()
    L.aq_xml_resolve.argtypes = [ctypes.c_char_p, ctypes.c_int,
                                 ctypes.c_char_p, ctypes.c_size_t,
                                 ctypes.POINTER(ctypes.c_size_t)]
    path = os.fsencode(os.path.abspath(xml_path))
    need = ctypes.c_size_t(0)
    _lib.check(L.aq_xml_resolve(path, dims, None, 0, ctypes.byref(need)))
    buf = ctypes.create_string_buffer(need.value)
    _lib.check(L.aq_xml_resolve(path, dims, buf, need.value,
                                ctypes.byref(need)))
    return [tuple(line.split("\t")[:3 if keys else 2]) for line in
            buf.value.decode().splitlines() if not line.startswith("#")]


def registered_kernels():
    """Registry keys of every kernel script this library implements."""
    L = _lib.lib()
    L.aq_sim_registered_kernel.restype = ctypes.c_char_p
    return [L.aq_sim_registered_kernel(i).decode()
            for i in range(L.aq_sim_n_registered_kernels())]


class Simulation:
    def __init__(self, xml_path, dims=3):
        L = _lib.lib()
        self._L = L
        L.aq_sim_create.argtypes = [ctypes.c_char_p, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_void_p)]
        L.aq_sim_destroy.argtypes = [ctypes.c_void_p]
        L.aq_sim_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.aq_sim_step_graph.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.aq_sim_sync.argtypes = [ctypes.c_void_p]
        L.aq_sim_stream.argtypes = [ctypes.c_void_p]
        L.aq_sim_stream.restype = ctypes.c_void_p
        L.aq_sim_n_particles.argtypes = [ctypes.c_void_p]
        L.aq_sim_n_particles.restype = ctypes.c_uint32
        L.aq_sim_var_info.argtypes = [ctypes.c_void_p, ctypes.c_char_p,
                                      ctypes.c_char_p, ctypes.c_size_t,
                                      ctypes.POINTER(ctypes.c_size_t),
                                      ctypes.POINTER(ctypes.c_size_t)]
        L.aq_sim_get.argtypes = [ctypes.c_void_p, ctypes.c_char_p,
                                 ctypes.c_void_p, ctypes.c_size_t]
        L.aq_sim_set.argtypes = [ctypes.c_void_p, ctypes.c_char_p,
                                 ctypes.c_void_p, ctypes.c_size_t]
        L.aq_sim_device_ptr.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        L.aq_sim_device_ptr.restype = ctypes.c_void_p
        L.aq_sim_n_tools.argtypes = [ctypes.c_void_p]
        L.aq_sim_tool_name.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.aq_sim_tool_name.restype = ctypes.c_char_p
        L.aq_sim_tool_type.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.aq_sim_tool_type.restype = ctypes.c_char_p
        L.aq_sim_definition.argtypes = [ctypes.c_void_p, ctypes.c_char_p,
                                        ctypes.c_char_p, ctypes.c_size_t]
        L.aq_sim_profile_tools.argtypes = [ctypes.c_void_p, ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_double)]
        L.aq_sim_walk_ms.argtypes = [ctypes.c_void_p,
                                     ctypes.POINTER(ctypes.c_double)]
        L.aq_sim_run.argtypes = [ctypes.c_void_p, ctypes.c_long, ctypes.c_int,
                                 ctypes.POINTER(ctypes.c_long)]
        L.aq_sim_save.argtypes = [ctypes.c_void_p]
        h = ctypes.c_void_p()
        _lib.check(L.aq_sim_create(os.fsencode(os.path.abspath(xml_path)),
                                   dims, ctypes.byref(h)))
        self._h = h
        self.dims = dims

    def close(self):
        if getattr(self, "_h", None):
            self._L.aq_sim_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- stepping -------------------------------------------------------
    def step(self, n=1, graph=False):
        f = self._L.aq_sim_step_graph if graph else self._L.aq_sim_step
        _lib.check(f(self._h, int(n)))

    def run(self, max_steps=0, graph=True):
        """The AQUAgpusph main loop: step until the case's Timing end
        criteria (at most max_steps steps when > 0), writing the particle
        <Save> files and AQUAgpusph.save.N.xml in the working directory at
        every output point.  Returns the steps run."""
        n = ctypes.c_long()
        _lib.check(self._L.aq_sim_run(self._h, int(max_steps), int(graph),
                                      ctypes.byref(n)))
        return n.value

    def save(self):
        """Write the <Save> files and the state file now."""
        _lib.check(self._L.aq_sim_save(self._h))

    def sync(self):
        _lib.check(self._L.aq_sim_sync(self._h))

    @property
    def stream(self):
        return self._L.aq_sim_stream(self._h)

    @property
    def N(self):
        return int(self._L.aq_sim_n_particles(self._h))

    # ---- variables ------------------------------------------------------
    def info(self, name):
        tb = ctypes.create_string_buffer(64)
        ln = ctypes.c_size_t()
        eb = ctypes.c_size_t()
        _lib.check(self._L.aq_sim_var_info(self._h, name.encode(), tb, 64,
                                           ctypes.byref(ln), ctypes.byref(eb)))
        return tb.value.decode(), ln.value, eb.value

    def get(self, name):
        t, n, eb = self.info(name)
        dt = _NP[_base(t)]
        comps = eb // np.dtype(dt).itemsize
        a = np.empty(n * comps, dt)
        _lib.check(self._L.aq_sim_get(self._h, name.encode(),
                                      a.ctypes.data_as(ctypes.c_void_p),
                                      a.nbytes))
        if t.endswith("*"):
            return a.reshape(n, comps) if comps > 1 else a
        return a if comps > 1 else a[0]

    def set(self, name, values):
        t, n, eb = self.info(name)
        a = np.ascontiguousarray(values, dtype=_NP[_base(t)])
        _lib.check(self._L.aq_sim_set(self._h, name.encode(),
                                      a.ctypes.data_as(ctypes.c_void_p),
                                      a.nbytes))

    def autotune(self, steps=3):
        """Time the fused-walk launch configurations on this case (the
        state is restored afterwards), keep the fastest; returns the
        report lines."""
        L = self._L
        L.aq_sim_autotune.argtypes = [ctypes.c_void_p, ctypes.c_int,
                                      ctypes.c_char_p, ctypes.c_size_t]
        buf = ctypes.create_string_buffer(4096)
        _lib.check(L.aq_sim_autotune(self._h, steps, buf, 4096))
        return buf.value.decode().splitlines()

    def device_ptr(self, name):
        return self._L.aq_sim_device_ptr(self._h, name.encode())

    def definition(self, name):
        b = ctypes.create_string_buffer(256)
        _lib.check(self._L.aq_sim_definition(self._h, name.encode(), b, 256))
        return b.value.decode()

    def tools(self):
        n = self._L.aq_sim_n_tools(self._h)
        return [(self._L.aq_sim_tool_name(self._h, i).decode(),
                 self._L.aq_sim_tool_type(self._h, i).decode())
                for i in range(n)]

    def profile(self, nsteps):
        """Mean GPU ms per tool over nsteps eager steps (in tool order)."""
        names = self.tools()
        ms = (ctypes.c_double * len(names))()
        _lib.check(self._L.aq_sim_profile_tools(self._h, int(nsteps), ms))
        return [(n, t, ms[i]) for i, (n, t) in enumerate(names)]

    def walk_ms(self):
        """Mean GPU ms of the fused neighbour-walk kernel alone over the last
        profile() call (None without a fused walk)."""
        v = ctypes.c_double()
        _lib.check(self._L.aq_sim_walk_ms(self._h, ctypes.byref(v)))
        return v.value if v.value >= 0 else None
