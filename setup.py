"""pip packaging of the engine, so that ``import libMHCUDA`` works without a
path edit, as with the reference's ``pip install`` (setup.py:62-65 there).

    python -m minhashcuda_amd.build          # HIP kernels + runtime + module
    pip install --no-build-isolation .       # installs libMHCUDA + minhashcuda_amd

The extension is the shared library that minhashcuda_amd/build.py compiles
with hipcc for gfx950 (``minhashcuda_amd/libMHCUDA.so``, built here when it is
missing); it is installed once, as the top-level module ``libMHCUDA``, which
the package's ctypes C ABI (minhashcuda_amd.LIB_PATH) then loads as well.
"""
import os
import shutil

from setuptools import Extension, setup
from setuptools.command.build_ext import build_ext

ROOT = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(ROOT, "minhashcuda_amd", "libMHCUDA.so")


class PrebuiltHip(build_ext):
    """Places the hipcc-built library where setuptools expects the extension,
    building it first if needed."""

    def run(self):
        if not os.path.exists(LIB):
            import sys
            sys.path.insert(0, ROOT)
            from minhashcuda_amd import build as hip_build
            hip_build.build()
        for ext in self.extensions:
            dest = self.get_ext_fullpath(ext.name)
            os.makedirs(os.path.dirname(dest) or ".", exist_ok=True)
            shutil.copyfile(LIB, dest)


setup(
    name="libMHCUDA",
    version="2.0.0",
    description="Weighted MinHash on AMD MI355X (gfx950), drop-in for src-d/minhashcuda's "
                "hashing path",
    packages=["minhashcuda_amd"],
    package_data={"minhashcuda_amd": ["csrc/*.h", "csrc/*.hip", "csrc/*.cpp"]},
    ext_modules=[Extension("libMHCUDA", sources=[])],
    cmdclass={"build_ext": PrebuiltHip},
    python_requires=">=3.8",
    install_requires=["numpy", "scipy"],
    zip_safe=False,
)
