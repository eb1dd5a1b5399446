"""CPU test of the pip packaging (setup.py): installed into a fresh target
directory, ``import libMHCUDA`` works without a path edit -- as after the
reference's ``pip install`` -- and the package's ctypes C ABI loads the same
single copy of the library as the module (no GPU is needed to import)."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_pip_install_gives_import_libMHCUDA(tmp_path):
    if not os.path.exists(os.path.join(ROOT, "minhashcuda_amd", "libMHCUDA.so")):
        pytest.skip("libMHCUDA.so is not built")
    # a copy of what setup.py packages, so that the build leaves nothing here
    src = tmp_path / "src"
    shutil.copytree(os.path.join(ROOT, "minhashcuda_amd"), src / "minhashcuda_amd",
                    ignore=shutil.ignore_patterns("_build", "__pycache__"))
    shutil.copy(os.path.join(ROOT, "setup.py"), src / "setup.py")
    target = tmp_path / "site"
    res = subprocess.run([sys.executable, "-m", "pip", "install", "--no-build-isolation",
                          "--no-deps", "--no-index", "--target", str(target), str(src)],
                         capture_output=True, text=True, timeout=600)
    assert res.returncode == 0, res.stdout + res.stderr
    probe = (
        "import libMHCUDA, minhashcuda_amd, os\n"
        "from minhashcuda_amd import capi\n"
        "assert callable(libMHCUDA.minhash_cuda_init) and callable(libMHCUDA.minhash_cuda_calc)\n"
        "assert minhashcuda_amd.LIB_PATH == libMHCUDA.__file__, minhashcuda_amd.LIB_PATH\n"
        "lib = capi.load(minhashcuda_amd.LIB_PATH)\n"
        "assert all(hasattr(lib, name) for name in capi.EXPORTS)\n"
        "print('ok')\n")
    env = dict(os.environ, PYTHONPATH=str(target))
    out = subprocess.run([sys.executable, "-c", probe], capture_output=True, text=True,
                         cwd=str(tmp_path), env=env, timeout=120)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout + out.stderr
