"""Loaders used by the tests: the product package (directory dpcuda-v2_amd, not an
importable name) and the CPU oracle backend (test infrastructure only)."""
import importlib.util
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "dpcuda-v2_amd")
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")


def _load(name, path):
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def pkg():
    return _load("dpcuda_v2_amd", os.path.join(PKG_DIR, "__init__.py"))


def workloads():
    return _load("dpcuda_v2_amd_workloads", os.path.join(PKG_DIR, "workloads.py"))


_oracle = None


def oracle():
    """The oracle's orc_ library wrapped by the same ctypes marshalling as libdpm."""
    global _oracle
    if _oracle is None:
        _oracle = pkg().Backend(ORACLE_SO, "orc_", gpu=False)
    return _oracle


def gpu_available():
    try:
        return pkg().device_count() > 0
    except Exception:
        return False
