"""How far one-ULP differences in the uniaxial transport's transcendentals
carry.

The uniaxial arms read acos, sin, cos, atan and atan2 (theta_c, the
effective extraordinary index, the walk-off, the complex square root of the
boundary solve).  The device's functions are not glibc's: they differ by an
ULP here and there, so the device cannot match the oracle path for path in
every case.  This measures, on the CPU, what one-ULP differences do: the
oracle built with those results moved by one ULP (libixo_jitter.so,
-DIXO_ULP_JITTER, about half the calls) against the oracle itself.  A
wave-normal direction near a channel's critical angle or near the exit
split's direction-match tolerance flips a decision, so a fraction of paths
takes another branch -- the scale the GPU parity bounds of uniaxial scenes
must allow (tests/test_film_gpu.py, tests/test_frosted_gpu.py).  Isotropic
and biaxial stones read none of these functions in their transport (the
biaxial indicatrix's general branch is algebraic) and are untouched.
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_lib
from indicatrix_amd import abi, tracer, workloads
from indicatrix_amd.tracer import events_to_numpy

JITTER = os.path.join(oracle_lib.ORACLE_DIR, "_build", "libixo_jitter.so")


def _paths(L, sc, spp):
    n = sc.width * sc.height * spp
    xyz = np.zeros((n, 3), np.float32)
    ev = (abi.PathEvents * n)()
    L.ixo_trace_paths(sc.ref(), 0, spp, 0, n, xyz.ctypes.data_as(C.POINTER(C.c_float)), ev, oracle_lib.threads())
    return xyz, events_to_numpy(ev)


def _flip(sc, spp=8):
    if not os.path.exists(JITTER):
        oracle_lib.build()
    J = C.CDLL(JITTER)
    J.ixo_trace_paths.argtypes = oracle_lib.lib().ixo_trace_paths.argtypes
    a, ea = _paths(oracle_lib.lib(), sc, spp)
    b, eb = _paths(J, sc, spp)
    same = (ea["bounces"] == eb["bounces"]) & (ea["termination"] == eb["termination"]) & \
           (ea["facet_hash"] == eb["facet_hash"])
    return 1.0 - same.mean(), a, b


def _top_view(material, preset="RingLights"):
    cam = tracer.camera_new(0.35, 1.15, 5.0, 18.0, 32, 24, 8)
    env = tracer.LightingPreset.studio(preset, 1.0, 0.85, 0.95)
    return tracer.FrameScene(cam, 32, 24, tracer.standard_round_brilliant(), material, 12, env)


def test_uniaxial_paths_move_with_one_ulp_transcendentals():
    flip_top, _, _ = _flip(_top_view(workloads.c2_material(bands=True)))
    flip_oval, _, _ = _flip(workloads.film_c2_scene(40, 30, 8, max_bounces=24))
    print(f"flip rate at one-ULP jitter: brilliant top view {flip_top:.2e}, C2 oval {flip_oval:.2e}")
    assert flip_top > 5e-3 and flip_oval > 1e-3


@pytest.mark.parametrize("material", ["diamond", "alexandrite"])
def test_isotropic_and_biaxial_paths_are_untouched(material):
    m = tracer.GemMaterial.diamond() if material == "diamond" else tracer.GemMaterial.alexandrite()
    flip, a, b = _flip(_top_view(m))
    assert flip == 0.0 and a.tobytes() == b.tobytes()
