"""The hand-off kernel's per-SIMD admission under concurrent launches.

k_trace_resume16 admits at most ceil(saved rays / (SIMDs x rays per
wave)) working waves per SIMD, counted in the launch's own control words,
and turns the other waves away.  Two contexts here launch plain jobs on
two streams back to back without waiting, so their hand-off kernels run at
the same time, each counting in its own context's words: a 4096-ray
shoebox (every ray alive at the hand-off, one working wave per SIMD), a
city block of 16 x 512 rays (a few waves per SIMD) and the city block of
16 x 4096 rays (every wave admitted).  Every launch must continue every
saved ray: per-ray records bit-identical to the job's synchronous
omg_trace_world, echograms within 1e-4 relative (the same f32
contributions summed in another atomic order).
"""
import numpy as np
import pytest
import torch   # before the HIP library is loaded: torch brings its own HIP runtime

import omg_hip as H
from omg_hip import scenes

pytestmark = pytest.mark.gpu

ECHO_TOL = 1e-4


def _jobs(c):
    mb, src, lis = scenes.c2_workload()
    city = H.Scene.from_builder(c, mb)
    mb, s1, l1, _ = scenes.shoebox(seed=1)
    box = H.Scene.from_builder(c, mb)
    return [("city block 16 x 4096 rays x 8 listeners", city,
             H.trace_params(4096, 3, seed=0xC2, bin_dt=0.001, max_time=2.0), src, lis),
            ("shoebox 1 x 4096 rays", box,
             H.trace_params(4096, 3, max_bounces=64, seed=0xC1, bin_dt=0.001, max_time=2.0,
                            floor_rel=1e-6), np.asarray([s1], np.float32), np.asarray([l1], np.float32)),
            ("city block 16 x 512 rays x 8 listeners", city,
             H.trace_params(512, 3, seed=7, bin_dt=0.001, max_time=2.0), src, lis)]


def _buffers(params, src, lis):
    dev = torch.device("cuda", 0)
    d_src = torch.from_numpy(np.ascontiguousarray(src, np.float32)).to(dev)
    d_lis = torch.from_numpy(np.ascontiguousarray(lis, np.float32)).to(dev)
    nb = H.nbins(params.max_time, params.bin_dt)
    n_src, n_lis = d_src.shape[0], d_lis.shape[0]
    # NaN-filled: each launch must write its echograms whole
    bins = torch.full((n_src, n_lis, nb, params.nbands), float("nan"), device=dev)
    dirs = torch.full((n_src, n_lis, nb, 3), float("nan"), device=dev)
    recs = torch.zeros((n_src * params.n_rays, 4), dtype=torch.int32, device=dev)
    return d_src, d_lis, bins, dirs, recs


def _launch(scene, params, bufs, stream):
    d_src, d_lis, bins, dirs, recs = bufs
    H.trace_world_async(scene, params, d_src.data_ptr(), d_src.shape[0], d_lis.data_ptr(),
                        d_lis.shape[0], bins.data_ptr(), dirs.data_ptr(), stream=stream.cuda_stream,
                        d_records=recs.data_ptr())


def _rel(a, b):
    scale = float(np.max(np.abs(b))) or 1.0
    return float(np.max(np.abs(a - b))) / scale


def test_concurrent_handoff_launches_continue_every_saved_ray(ctx):
    ctx.set_launch_config()
    ctx.set_handoff()
    ctx.set_handoff_lanes()
    ctx.set_route_rays()
    other = H.Context(0)
    jobs = []
    try:
        jobs = [_jobs(ctx), _jobs(other)]
        refs = []
        for name, scene, params, src, lis in jobs[0]:
            r = H.trace_world(scene, params, src, lis, records=True)
            refs.append((name, r.records.reshape(-1, 4), r.bins, r.dirs))
        streams = [torch.cuda.Stream(device=torch.device("cuda", 0)) for _ in range(2)]
        outs = []
        for rep in range(2):
            for j in range(len(refs)):
                for side in range(2):
                    _, scene, params, src, lis = jobs[side][j]
                    outs.append((j, side, rep, scene, params, _buffers(params, src, lis)))
        torch.cuda.synchronize()
        # every launch back to back on the two streams, no waits in between
        for j, side, rep, scene, params, bufs in outs:
            _launch(scene, params, bufs, streams[side])
        torch.cuda.synchronize()
        for j, side, rep, _, _, (_, _, bins, dirs, recs) in outs:
            name, rec_ref, bins_ref, dirs_ref = refs[j]
            rec = recs.cpu().numpy().view(np.uint32)
            diff = np.nonzero(np.any(rec != rec_ref, axis=1))[0]
            assert diff.size == 0, (f"{name}, context {side}, pass {rep}: {diff.size} rays differ, "
                                    f"first {diff[:5].tolist()}")
            eb = _rel(bins.cpu().numpy(), bins_ref)
            ed = _rel(dirs.cpu().numpy(), dirs_ref)
            print(f"{name}, context {side}, pass {rep}: records identical, echogram rel diff "
                  f"{eb:.1e} / dirs {ed:.1e}", flush=True)
            assert eb <= ECHO_TOL and ed <= ECHO_TOL
    finally:
        for side in jobs:
            for _, scene, _, _, _ in side:
                scene.close()
        other.close()
