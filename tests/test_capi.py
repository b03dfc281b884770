"""The C-ABI library loads on a host without a GPU and exports every entry point
declared in include/qmcx.h; the Python host layer binds exactly those.  No compute
call is made (there is no device here); argument checks that fail before any
device work are exercised."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from miniqmc_amd import qmcx, systems

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "qmcx.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qmcx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "qmcx_run_sweeps" in names and "qmcx_batch_create" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    assert os.path.exists(qmcx.LIB_PATH), "libqmcx.so not built"
    out = subprocess.run(["nm", "-D", "--defined-only", qmcx.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (qmcx_[a-z0-9_]+)", out))
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing


def test_python_binding_covers_header():
    assert set(qmcx.exported_symbols()) == set(declared())


def test_library_loads_without_gpu():
    L = qmcx.load()
    assert isinstance(L.qmcx_last_error(), bytes)


def test_argument_errors_before_device_work():
    L = qmcx.load()
    h = C.c_void_p()
    lat = np.eye(3).reshape(-1)
    ions = np.zeros((2, 3)).reshape(-1)
    # nup != ndn is rejected before any HIP call
    rc = L.qmcx_system_create(C.byref(h), lat, 2, ions, np.array([0, 1], dtype=np.int32), 2, 4, 5)
    assert rc == -1 and b"nup" in L.qmcx_last_error()
    rc = L.qmcx_system_create(C.byref(h), lat, 2, ions, np.array([0, 7], dtype=np.int32), 2, 4, 4)
    assert rc == -1 and b"species" in L.qmcx_last_error()
    rc = L.qmcx_invert_transpose(0, 1, np.zeros(1), np.zeros(1), np.zeros(1), np.zeros(1))
    assert rc == -1


def test_systems_generate_reference_shapes():
    s = systems.make_system("nio32")
    assert (s.nel, s.norb, s.npad, s.ng) == (384, 192, 208, (48, 48, 48))
    assert s.table_bytes() == 51 ** 3 * 208 * 4          # 110,365,632 B
    assert s.nions == 32 and (s.ion_species == 0).sum() == 16
    c = s.coefs.reshape(51, 51, 51, 208)
    assert np.array_equal(c[48:], c[:3]) and np.all(c[..., 192:] == 0)
    assert c[..., :192].min() >= -1 and c[..., :192].max() < 1
    t = systems.make_system("nio16")
    assert (t.nel, t.norb, t.nions) == (192, 96, 16)
    # NiO-64 ions come from tiling (1,1,2), grouped by species (expandSuperCell)
    sup, pos, spc = systems.tile_ions(systems.NIO_CELL, systems.NIO_FRAC, systems.NIO_SPECIES, (1, 1, 2))
    assert len(pos) == 64 and np.all(np.diff(spc) >= 0)
    assert abs(sup[2, 2] - 2 * 15.7622) < 1e-12
