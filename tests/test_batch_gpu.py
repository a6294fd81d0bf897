"""ixg_accumulate_batch against each picture traced alone.

Mirrors the reference's batch tests (renderer/gpu_backend/tests.rs:497-860):
a batch of different pictures equals each picture traced alone bit for bit
and is reported once per picture in input order; a cancel raised from the
first on_done finishes what was queued and cancels the rest with their
outputs untouched; a device loss declines every unreported picture and the
backend recovers; an unsupported material declines only its own picture.
"""
import ctypes as C

import numpy as np
import pytest

from indicatrix_amd import designs, tracer, workloads

pytestmark = pytest.mark.gpu

DONE = tracer.GpuAccumulate.DONE
DECLINED = tracer.GpuAccumulate.DECLINED
CANCELLED = tracer.GpuAccumulate.CANCELLED


def _scene(material, yaw, size):
    cam = tracer.camera_new(yaw, 0.28, 5.0, 18.0, size[0], size[1], 1)
    env = tracer.LightingPreset.studio("RingLights", 1.0, 0.85, 0.95)
    return tracer.FrameScene(cam, size[0], size[1], tracer.standard_round_brilliant(), material, 16, env)


def _seeded_material():
    disp, _, _ = designs.seeded_material(7)
    m = tracer.GemMaterial.diamond().to_abi()
    m.dispersion = disp
    return m


def three_different_pictures():
    """Different material class (isotropic, seeded isotropic, uniaxial
    pleochroic), size, camera and sample range (tests.rs:546-566)."""
    return [(_scene(tracer.GemMaterial.diamond(), 0.35, (64, 48)), 0, 3),
            (_scene(_seeded_material(), 0.9, (40, 40)), 5, 2),
            (_scene(workloads.c2_material(), 1.7, (56, 32)), 0, 4)]


def traced_alone(backend, specs):
    outs = []
    for scene, first, samples in specs:
        acc = np.zeros((scene.width * scene.height, 3))
        assert backend.try_accumulate_cancellable(scene, first, samples, acc) is DONE
        outs.append(acc)
    return outs


def traced_as_batch(backend, specs, cancel=None, on_done=None):
    outs = [np.zeros((s.width * s.height, 3)) for s, _, _ in specs]
    outcomes = backend.try_accumulate_batch_cancellable(
        [(s, f, n, o) for (s, f, n), o in zip(specs, outs)], cancel, on_done)
    return outs, outcomes


def test_a_batch_of_different_scenes_equals_each_scene_alone(backend):
    specs = three_different_pictures()
    print("batch: 3 pictures vs alone", flush=True)
    expected = traced_alone(backend, specs)
    order = []
    outs, outcomes = traced_as_batch(backend, specs, on_done=lambda i, _: order.append(i))
    assert order == [0, 1, 2]
    assert outcomes == [DONE, DONE, DONE]
    for i in range(3):
        assert outs[i].tobytes() == expected[i].tobytes(), i
    assert any(o.sum() > 0 for o in outs)


def test_multi_chunk_pictures_stay_bit_identical():
    """Small chunks: every picture spans several launches, and the next
    picture's chunks are queued before the previous one is read back."""
    b = tracer.GpuBackend(chunk_paths=2048)
    try:
        specs = three_different_pictures()
        print("batch: multi-chunk", flush=True)
        expected = traced_alone(b, specs)
        outs, outcomes = traced_as_batch(b, specs * 2)
        assert outcomes == [DONE] * 6
        for i in range(6):
            assert outs[i].tobytes() == expected[i % 3].tobytes(), i
    finally:
        b.close()


def test_cancelling_mid_batch_reports_cancelled_and_leaves_the_backend_reusable(backend):
    specs = three_different_pictures()
    expected = traced_alone(backend, specs)
    flag = (C.c_uint8 * 1)(0)

    def on_done(i, outcome):
        if i == 0 and outcome is DONE:
            flag[0] = 1

    print("batch: cancel from on_done", flush=True)
    outs, outcomes = traced_as_batch(backend, specs, cancel=flag, on_done=on_done)
    assert outcomes[0] is DONE and outs[0].tobytes() == expected[0].tobytes()
    assert outcomes[1] in (DONE, CANCELLED)   # was already queued when the flag fired
    assert outcomes[2] is CANCELLED
    for i, o in enumerate(outcomes):
        if o is DONE:
            assert outs[i].tobytes() == expected[i].tobytes()
        else:
            assert not outs[i].any(), f"cancelled picture {i} must leave its output untouched"
    assert not backend.is_lost()
    outs, outcomes = traced_as_batch(backend, specs)
    assert outcomes == [DONE] * 3
    assert all(outs[i].tobytes() == expected[i].tobytes() for i in range(3))


def test_a_cancel_before_the_batch_cancels_everything(backend):
    specs = three_different_pictures()
    flag = (C.c_uint8 * 1)(1)
    outs, outcomes = traced_as_batch(backend, specs, cancel=flag)
    assert outcomes == [CANCELLED] * 3
    assert not any(o.any() for o in outs)


def test_a_device_loss_declines_the_rest_and_the_backend_recovers():
    b = tracer.GpuBackend()
    try:
        specs = three_different_pictures()
        expected = traced_alone(b, specs)
        b.inject_fault(0)   # the first chunk of the batch fails
        print("batch: injected device loss", flush=True)
        outs, outcomes = traced_as_batch(b, specs)
        assert outcomes == [DECLINED] * 3
        assert not any(o.any() for o in outs)
        assert b.is_lost()
        outs, outcomes = traced_as_batch(b, specs)   # still lost: everything declines
        assert outcomes == [DECLINED] * 3
        assert b.try_recover()
        outs, outcomes = traced_as_batch(b, specs)
        assert outcomes == [DONE] * 3
        assert all(outs[i].tobytes() == expected[i].tobytes() for i in range(3))
    finally:
        b.close()


def test_an_unsupported_material_declines_only_its_picture(backend):
    specs = three_different_pictures()
    expected = traced_alone(backend, specs)
    biaxial = tracer.GemMaterial.diamond().to_abi()
    biaxial.biaxial = 1
    bad = (_scene(biaxial, 0.5, (16, 16)), 0, 2)
    print("batch: declined middle picture", flush=True)
    outs, outcomes = traced_as_batch(backend, [specs[0], bad, specs[2]])
    assert outcomes == [DONE, DECLINED, DONE]
    assert outs[0].tobytes() == expected[0].tobytes()
    assert not outs[1].any()
    assert outs[2].tobytes() == expected[2].tobytes()


def test_empty_batch_and_zero_samples(backend):
    assert backend.try_accumulate_batch_cancellable([]) == []
    scene, _, _ = three_different_pictures()[0]
    out = np.full((scene.width * scene.height, 3), 2.0)
    assert backend.try_accumulate_batch_cancellable([(scene, 0, 0, out)]) == [DONE]
    assert np.all(out == 2.0)
