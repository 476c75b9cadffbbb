"""Writes the reference's own tool-test cases as fixtures under
tests/golden/refcases/: the case file main.xml of tests/{2D,3D}/<Tool>/cMake
and of tests/ExternalTool/cMake, with the XML comments (the licence banner)
dropped and nothing else changed (the @RESOURCES_DIR@ and
@EXTTOOL_LIB_NAME@ placeholders stay: the GPU test substitutes them, as the
reference's CMake configure step does), plus the particles.dat each case
loads, stored once per distinct file.  manifest.json maps each case to its
particle file.

Run in a container that holds the reference checkout:
    python tests/golden/make_refcases.py [/root/reference]

Only cases whose tools are all native types are taken: Assert, Copy,
RadixSort, Reduction, Set, SetScalar, Sort (2D and 3D) and the installable
ExternalTool demo.  Kernel, LinkList and MLS load OpenCL kernels of their
own (init.cl, grad_p.cl), which this library does not compile.
"""
import hashlib
import json
import os
import re
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "refcases")
TOOLS = ("Assert", "Copy", "RadixSort", "Reduction", "Set", "SetScalar",
         "Sort")


def main(ref):
    os.makedirs(OUT, exist_ok=True)
    known = {}
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".dat"):
            with open(os.path.join(HERE, f), "rb") as fh:
                known.setdefault(hashlib.md5(fh.read()).hexdigest(), f)
    cases = [("%s_%s" % (d, t), os.path.join(ref, "tests", d, t, "cMake"))
             for d in ("2D", "3D") for t in TOOLS]
    cases.append(("ExternalTool",
                  os.path.join(ref, "tests", "ExternalTool", "cMake")))
    manifest = {}
    for name, src in cases:
        with open(os.path.join(src, "main.xml")) as fh:
            xml = fh.read()
        xml = re.sub(r"<!--.*?-->\s*", "", xml, flags=re.S)
        with open(os.path.join(OUT, name + ".xml"), "w") as fh:
            fh.write(xml)
        with open(os.path.join(src, "particles.dat"), "rb") as fh:
            data = fh.read()
        key = hashlib.md5(data).hexdigest()
        if key not in known:
            fname = "refcase_%s_particles.dat" % name.lower()
            with open(os.path.join(HERE, fname), "wb") as fh:
                fh.write(data)
            known[key] = fname
        manifest[name] = {"particles": known[key],
                          "dims": 3 if name.startswith("3D") else 2}
    with open(os.path.join(OUT, "manifest.json"), "w") as fh:
        json.dump(manifest, fh, indent=1, sort_keys=True)
        fh.write("\n")
    print("wrote %d cases to %s" % (len(manifest), OUT))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
