"""Builds gy_autograd (the LibTorch autograd.Function over libgyunity.so) in
place: gyunity_amd/gy_autograd*.so.  Plain C++ against torch's headers; the
GPU work happens in libgyunity.so.

  python gyunity_amd/torch_ext/build.py
"""
import os
import sys

from setuptools import setup
from torch.utils.cpp_extension import BuildExtension, CppExtension

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)


def main():
    ext = CppExtension(
        name="gy_autograd",
        sources=[os.path.join(HERE, "gy_autograd.cpp")],
        include_dirs=[os.path.join(ROOT, "include")],
        library_dirs=[PKG],
        libraries=["gyunity"],
        extra_compile_args=["-O2", "-std=c++17", "-g0"],
        extra_link_args=["-Wl,-rpath,$ORIGIN", "-s"],
    )
    cwd = os.getcwd()
    os.chdir(PKG)
    try:
        setup(name="gy_autograd", ext_modules=[ext], cmdclass={"build_ext": BuildExtension},
              script_args=["build_ext", "--inplace", "--build-temp", os.path.join(PKG, "_build", "torch_ext"),
                           "--build-lib", os.path.join(PKG, "_build", "torch_lib")])
    finally:
        os.chdir(cwd)


if __name__ == "__main__":
    sys.exit(main())
