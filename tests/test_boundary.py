"""The C-ABI library: it loads, exports every symbol declared in include/*.h,
and fails loudly (no CPU fallback) when no GPU is present."""
import ctypes
import glob
import os
import re

import pytest

from tests.conftest import ROOT, gpu_available


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b([A-Za-z_]\w*)\s*\(", text):
            name = m.group(1)
            prefix = text[:m.start()].rstrip()
            # function declarations are preceded by a return type
            if re.search(r"(int|void|char\s*\*|double|size_t|const char\s*\*|\*)\s*$", prefix):
                names.add(name)
    return names


def test_library_exports_every_declared_symbol():
    lib_path = os.path.join(ROOT, "pacific_amd", "libgrains_gpu.so")
    if not os.path.exists(lib_path):
        pytest.fail("libgrains_gpu.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(lib_path)
    names = declared_symbols()
    assert len(names) > 20
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_mirrors_the_abi():
    from pacific_amd import grains
    assert set(grains.EXPORTS) <= declared_symbols()


def test_no_cpu_fallback_without_gpu():
    if gpu_available():
        pytest.skip("GPU present")
    from pacific_amd import GrainsError, GrainsGPU
    g = GrainsGPU((0, 0, 0), (0.01, 0.01, 0.01))
    g.contact_model(0, 0, 1e4, 0.9, 0.5)
    g.add_spheres(1e-3, 2500.0, 5e-5, 0, [[0.005, 0.005, 0.005]])
    with pytest.raises(GrainsError) as e:
        g.init()
    assert e.value.kind == "DeviceError"


def test_argument_errors_are_reported():
    from pacific_amd import GrainsError, GrainsGPU
    g = GrainsGPU((0, 0, 0), (0.01, 0.01, 0.01))
    with pytest.raises(GrainsError):
        g.add_spheres(-1.0, 2500.0, 5e-5, 0, [[0.005, 0.005, 0.005]])
    with pytest.raises(GrainsError):
        g.contact_model(0, 0, 1e4, 1.5, 0.5)


def test_cartesian_plugin_shim_compiles_and_links(tmp_path):
    # include/GrainsCoupledWithFluid.hh (the C++ calls of FS_Grains3DPlugIn
    # over the C ABI) compiles as plain C++ and links against the library
    import shutil
    import subprocess
    cxx = shutil.which("g++")
    if not cxx:
        pytest.skip("no g++")
    src = tmp_path / "plugin.cpp"
    src.write_text('#include "GrainsCoupledWithFluid.hh"\n'
                   "int main(int argc, char** argv) {\n"
                   "  if (argc < 2) return 0;\n"
                   "  GrainsCoupledWithFluid* g = GrainsCoupledWithFluid::create(argv[1], 1000.);\n"
                   "  g->setFluidCorrectedAcceleration(true);\n"
                   "  g->do_before_time_stepping();\n"
                   "  std::vector<std::vector<double> > ft(1, std::vector<double>(6, 0.));\n"
                   "  g->updateParticlesHydroFT(&ft);\n"
                   "  g->Simulation(1e-4);\n"
                   "  std::istringstream is;\n"
                   "  g->GrainsToFluid(is);\n"
                   "  g->doPostProcessing();\n"
                   "  delete g;\n"
                   "  return 0;\n"
                   "}\n")
    lib_dir = os.path.join(ROOT, "pacific_amd")
    r = subprocess.run([cxx, "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src),
                        "-L", lib_dir, "-lgrains_gpu", "-Wl,--allow-shlib-undefined", "-o", str(tmp_path / "plugin")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
