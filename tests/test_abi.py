"""The C-ABI library loads and exports every symbol include/tvk_hip.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

from tuvok_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "tvk_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(tvk_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_the_python_view():
    assert set(declared()) == set(_abi.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_abi.LIB_PATH)
    missing = [s for s in declared() if not hasattr(lib, s)]
    assert not missing, missing


def test_default_params_are_the_reference_defaults():
    L = _abi.load()
    p = _abi.Params()
    assert L.tvk_default_params(ctypes.byref(p)) == 0
    # AbstrRenderer.cpp:71-214: 1D TF, lighting on, srm 1, ambient 0.1, light dir (0,0,-1)
    assert p.mode == _abi.RM_1DTRANS and p.lighting == 1 and p.srm == 1.0
    assert abs(p.light_ambient[0] - 0.1) < 1e-7 and tuple(p.light_dir) == (0.0, 0.0, -1.0)
    assert p.refine_steps == 5 and p.ert_eps == 0.0


def test_null_arguments_are_rejected_without_a_gpu():
    L = _abi.load()
    assert L.tvk_create(None, 0) == -1
    assert L.tvk_default_params(None) == -1


def test_cpp_front_end_compiles(tmp_path):
    """include/tvk_raycaster.hpp (the AbstrRenderer-shaped C++ front end) compiles."""
    import subprocess
    src = tmp_path / "use.cpp"
    src.write_text('#include "tvk_raycaster.hpp"\n'
                   'int main() { tvk::Raycaster r(0); r.SetRendermode(TVK_RM_1DTRANS); r.Resize(8, 8);\n'
                   '  auto img = r.Paint(); return (int)img.size() > 0 ? 0 : 1; }\n')
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_batch_abi_defaults_and_null_arguments():
    L = _abi.load()
    t = _abi.Tuning()
    assert L.tvk_default_tuning(ctypes.byref(t)) == 0
    assert (t.tile_w, t.tile_h, t.groups_per_cu, t.use_graph) == (8, 8, 0, 1)
    assert L.tvk_batch_create(None, 0) == -1
    assert L.tvk_batch_render(None, None) == -1
    assert L.tvk_batch_add(None, None, None, 8, 8, None) == -1
    assert L.tvk_share_dataset(None, None) == -1
