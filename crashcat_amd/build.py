"""Builds the gfx950 shared library libcrashcat_amd.so in-tree with hipcc."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "crashcat_amd.hip")
OUT = os.path.join(HERE, "libcrashcat_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # no packed FP32 instructions in generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    # keep a*b+c as two roundings, as the FP64 oracle computes it
    "-ffp-contract=off",
]


def build(verbose=False):
    deps = [os.path.join(HERE, "csrc", f) for f in os.listdir(os.path.join(HERE, "csrc"))]
    deps.append(os.path.join(HERE, "..", "include", "crashcat_amd.h"))
    cmd = [HIPCC] + FLAGS + ["-o", OUT + ".tmp", SRC]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(OUT + ".tmp", OUT)
    return OUT


NODE_DIR = os.path.join(HERE, "node")
NODE_SRC = os.path.join(NODE_DIR, "cc_node.c")
NODE_OUT = os.path.join(NODE_DIR, "crashcat_amd.node")
NODE_INCLUDE = "/usr/include/node"


def node_available():
    """The Node N-API headers are installed (the addon is plain C over node_api.h)."""
    return os.path.exists(os.path.join(NODE_INCLUDE, "node_api.h"))


def build_node(verbose=False):
    """Builds the N-API addon crashcat_amd/node/crashcat_amd.node against the library."""
    cmd = ["gcc", "-O2", "-Wall", "-shared", "-fPIC", "-DNODE_GYP_MODULE_NAME=crashcat_amd", "-I" + NODE_INCLUDE,
           "-I" + os.path.join(HERE, "..", "include"), NODE_SRC, "-L" + HERE, "-lcrashcat_amd",
           "-Wl,-rpath,$ORIGIN/..", "-o", NODE_OUT + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(NODE_OUT + ".tmp", NODE_OUT)
    return NODE_OUT


if __name__ == "__main__":
    build(verbose=True)
    print("built", OUT)
    if node_available():
        build_node(verbose=True)
        print("built", NODE_OUT)
    sys.exit(0)
