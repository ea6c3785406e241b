"""Flat wire formats on the CPU: the web bridge's world-job layout decoded
as web/gpu.js reads it, and ParamBlock::write_flat / read_flat
(crates/omg-core/src/params.rs:62-143) against a restatement of the
reference's writer, offset by offset."""
import numpy as np
import pytest

import omg_hip as H
from omg_hip import wire


def world_job(sid, rays, seed, src, lis, panels):
    """The writer of crates/omg-web/src/lib.rs:394-410 (fixed 1034-f32 stride)."""
    j = [float(sid), float(min(rays * 8, 8192)), float(seed)]
    j += list(src) + list(lis)
    np_ = min(len(panels), 64)
    j.append(float(np_))
    for mn, mx, (ab, sc, tr) in panels[:np_]:
        j += [mn[0], mn[1], mn[2], sc, mx[0], mx[1], mx[2], 0.0]
        j += [ab[0], ab[1], ab[2], 0.0, tr[0], tr[1], tr[2], 0.0]
    j += [0.0] * (1034 - len(j))
    return np.array(j, np.float32)


def test_world_job_decodes_as_the_js_driver_reads_it():
    rng = np.random.default_rng(4)
    seed = 0xC6D2
    for k in range(20):
        seed = (seed + 0x9E3779B9) & 0xFFFFFFFF          # lib.rs:398-400
        panels = [((rng.uniform(-5, 0, 3)), rng.uniform(0, 5, 3),
                   (rng.uniform(0, 1, 3), rng.uniform(0, 1), rng.uniform(0, 1, 3)))
                  for _ in range(int(rng.integers(0, 70)))]
        src, lis = rng.uniform(-9, 9, 3), rng.uniform(-9, 9, 3)
        j = world_job(k + 3, int(rng.integers(1, 2000)), seed, src, lis, panels)
        d = wire.decode_world_job(j)
        assert d.id == k + 3
        assert d.n_rays == int(j[1])
        # Uint32Array assignment: ToUint32 of the f32 word (2^32 wraps to 0)
        assert d.seed == int(np.float64(j[2])) % (1 << 32)
        assert np.array_equal(np.array(d.src, np.float32), j[3:6])
        assert np.array_equal(np.array(d.lis, np.float32), j[6:9])
        assert d.n_panels == min(len(panels), 64)
        for i in range(d.n_panels):
            q = j[10 + 16 * i:26 + 16 * i]
            p = d.panels[i]
            assert np.array_equal(np.array(p.min[:], np.float32), q[0:3])
            assert np.array_equal(np.array(p.max[:], np.float32), q[4:7])
            assert np.float32(p.mat.scattering) == q[3]
            assert np.array_equal(np.array(p.mat.absorption[:3], np.float32), q[8:11])
            assert np.array_equal(np.array(p.mat.transmission[:3], np.float32), q[12:15])
    j = world_job(1, 10, 5, [0, 0, 0], [1, 1, 1], [])
    j[9] = 100.0                                           # Math.min(jobs[9], 64)
    assert wire.decode_world_job(j).n_panels == 64
    with pytest.raises(H.OmgError) as e:
        wire.decode_world_job(j, version=2)
    assert e.value.code == H._lib.ERR_VERSION


def write_flat_reference(p):
    """ParamBlock::write_flat (params.rs:82-107), offset by offset."""
    out = np.zeros(24 + 9 * len(p.taps), np.float32)
    out[0] = np.float32(p.version)
    out[1] = np.float32(len(p.taps))
    for b in range(3):
        out[2 + b] = p.rt60[b]
        out[5 + b] = p.level[b]
    if p.remote:
        out[8] = 1.0
        for b in range(3):
            out[9 + b] = p.remote["rt60"][b]
            out[12 + b] = p.remote["send"][b]
        out[15:24] = p.remote["sh"]
    for i, (key, delay, d, g) in enumerate(p.taps):
        o = 24 + 9 * i
        out[o] = np.float32(key)
        out[o + 1] = delay
        out[o + 2:o + 5] = d
        out[o + 5:o + 8] = g
    return out


def test_param_block_flat_layout_and_round_trip():
    rng = np.random.default_rng(9)
    for remote in (None, dict(rt60=list(rng.uniform(0.1, 5, 3)), send=list(rng.uniform(0, 1, 3)),
                              sh=list(rng.uniform(-1, 1, 9)))):
        taps = [(int(rng.integers(0, 1 << 24)), float(rng.uniform(0, 0.2)),
                 tuple(rng.uniform(-1, 1, 3)), tuple(rng.uniform(0, 1, 3))) for _ in range(17)]
        p = wire.Params(version=123456, rt60=tuple(rng.uniform(0.1, 5, 3)),
                        level=tuple(rng.uniform(0, 0.7, 3)), remote=remote, taps=taps)
        flat = wire.param_block_write_flat(p)
        assert np.array_equal(flat, write_flat_reference(p))
        q = wire.param_block_read_flat(flat)
        assert q.version == 123456 and len(q.taps) == 17
        assert np.array_equal(np.array(q.rt60, np.float32), np.array(p.rt60, np.float32))
        assert (q.remote is None) == (remote is None)
        assert np.array_equal(wire.param_block_write_flat(q), flat)
    with pytest.raises(H.OmgError):
        wire.param_block_read_flat(flat[:30])             # shorter than its taps
