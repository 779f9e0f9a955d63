"""The R package skeleton (gpusgdlm_amd/rpkg) is consistent with the
engine it links: package and module names as rSGDLM's, useDynLib of the
package, the Makevars pointing at this repository's C++ interface headers
and at the two libraries the build produces.  R itself is not installed
here, so the package build is not exercised."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "gpusgdlm_amd", "rpkg")


def _fields(path):
    out, key = {}, None
    for line in open(path):
        if line[:1].isspace() and key:
            out[key] += " " + line.strip()
        elif ":" in line:
            key, val = line.split(":", 1)
            out[key] = val.strip()
    return out


def test_description_names_the_module():
    d = _fields(os.path.join(PKG, "DESCRIPTION"))
    assert d["Package"] == "rSGDLM"
    assert d["RcppModules"] == "gpuSGDLM"
    assert "Rcpp" in d["LinkingTo"] and "Rcpp" in d["Imports"]


def test_namespace_loads_the_package_library():
    ns = open(os.path.join(PKG, "NAMESPACE")).read()
    assert re.search(r"useDynLib\(rSGDLM\b", ns)
    assert "importFrom(Rcpp" in ns


def test_makevars_links_the_engine():
    mk = open(os.path.join(PKG, "src", "Makevars")).read()
    assert "-lsgdlm_host" in mk and "-lsgdlm_hip" in mk
    assert "-DUSE_RCPP" in mk and "$(GPUSGDLM_HOME)/include" in mk
    objs = re.search(r"^OBJECTS\s*=\s*(.*)$", mk, re.M).group(1).split()
    assert "R_main.o" in objs and "SGDLM/RcppWrapper.o" in objs
    # the headers the Rcpp layer includes as SGDLM/... exist in include/
    for h in ("HostWrapper.hpp", "HostWrapperFactory.hpp"):
        assert os.path.isfile(os.path.join(ROOT, "include", "SGDLM", h))
    # and the build names its libraries as the Makevars links them
    build = open(os.path.join(ROOT, "gpusgdlm_amd", "build.py")).read()
    assert "libsgdlm_host" in build and "libsgdlm_hip" in build
