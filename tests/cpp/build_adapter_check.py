"""Build tests/cpp/build/adapter_check (the check program of include/agama_amd.hpp) with g++
against the in-tree libagama_amd.so: test infrastructure, run by __graft_entry__.build() and by
tests/test_cpp_adapter.py."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "build", "adapter_check")


def build():
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    libdir = os.path.join(ROOT, "agama_amd")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include"),
           os.path.join(HERE, "adapter_check.cpp"), "-o", OUT + ".tmp", "-L", libdir, "-lagama_amd",
           "-Wl,-rpath," + libdir + ",-rpath,$ORIGIN/../../../agama_amd"]
    subprocess.check_call(cmd)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build())
