"""The allocation-free launch (omg_trace_world_async_noalloc) through the C
ABI against omg_trace_world_async, the entry point whose kernels it
shares, on the same device buffers: records bit-identical, echograms within
1e-4 relative (the two sum the same f32 contributions in different atomic
orders).  Two calls in a row - the city block workload (16 sources x 4096
rays x 8 listeners), then a smaller job with an overlay panel, whose
tables differ - then the first job again; captured in a HIP graph and
replayed, it equals its eager call on every replay; the limits are refused.
"""
import numpy as np
import pytest
import torch   # before the HIP library is loaded: torch brings its own HIP runtime

import omg_hip as H
from omg_hip import scenes
from helpers import echogram_rel_errors

pytestmark = pytest.mark.gpu

ECHO_TOL = 1e-4


def _buffers(params, n_src, n_lis):
    dev = torch.device("cuda", 0)
    nb = H.nbins(params.max_time, params.bin_dt)
    # NaN-filled: the call itself must clear the echograms
    bins = torch.full((n_src, n_lis, nb, params.nbands), float("nan"), device=dev)
    dirs = torch.full((n_src, n_lis, nb, 3), float("nan"), device=dev)
    recs = torch.zeros((n_src * params.n_rays, 4), dtype=torch.int32, device=dev)
    return bins, dirs, recs


def _launch(fn, scene, params, d_src, d_lis, overlays, stream):
    bins, dirs, recs = _buffers(params, d_src.shape[0], d_lis.shape[0])
    torch.cuda.synchronize()
    fn(scene, params, d_src.data_ptr(), d_src.shape[0], d_lis.data_ptr(), d_lis.shape[0],
       bins.data_ptr(), dirs.data_ptr(), stream=stream.cuda_stream, overlays=overlays,
       d_records=recs.data_ptr())
    return bins, dirs, recs


def _host(out):
    torch.cuda.synchronize()
    bins, dirs, recs = out
    return recs.cpu().numpy().view(np.uint32), bins.cpu().numpy(), dirs.cpu().numpy()


def test_noalloc_matches_async_on_consecutive_calls(ctx):
    ctx.set_launch_config()
    ctx.set_handoff()
    ctx.set_handoff_lanes()
    ctx.set_route_rays()
    dev = torch.device("cuda", 0)
    mb, src, lis = scenes.c2_workload()
    scene = H.Scene.from_builder(ctx, mb)
    l0 = lis[0]
    panel = ((l0[0] + 1.0, l0[1] - 3.0, l0[2] - 1.0), (l0[0] + 1.2, l0[1] + 3.0, l0[2] + 2.0),
             H.MaterialDef([0.1, 0.2, 0.3], 0.2, [0.5, 0.4, 0.3]))
    jobs = [("city block 16 x 4096 rays x 8 listeners",
             H.trace_params(4096, 3, seed=0xC2, bin_dt=0.001, max_time=2.0), src, lis, ()),
            ("3 x 1000 rays x 2 listeners with a panel",
             H.trace_params(1000, 3, seed=77, bin_dt=0.001, max_time=1.0), src[[1, 9, 14]],
             lis[[0, 5]], (panel,))]
    jobs.append(jobs[0])
    stream = torch.cuda.Stream(device=dev)
    dev_in = [(torch.from_numpy(np.ascontiguousarray(s, np.float32)).to(dev),
               torch.from_numpy(np.ascontiguousarray(l, np.float32)).to(dev)) for _, _, s, l, _ in jobs]
    # the allocation-free calls back to back, then the reference entry point
    got = [_host(_launch(H.trace_world_async_noalloc, scene, p, *dev_in[k], ov, stream))
           for k, (_, p, _, _, ov) in enumerate(jobs)]
    ref = [_host(_launch(H.trace_world_async, scene, p, *dev_in[k], ov, stream))
           for k, (_, p, _, _, ov) in enumerate(jobs)]
    for k, (label, p, s, l, _) in enumerate(jobs):
        g_rec, g_bins, g_dirs = got[k]
        r_rec, r_bins, r_dirs = ref[k]
        print(f"case no-alloc call {k}: {label}", flush=True)
        assert not np.isnan(g_bins).any() and not np.isnan(g_dirs).any(), label
        diff = np.nonzero(np.any(g_rec != r_rec, axis=1))[0]
        assert diff.size == 0, f"{label}: {diff.size} rays differ, first {diff[:5]}"
        assert np.any(g_rec[:, 0] != 0), f"{label}: no records written"
        eb, ed = echogram_rel_errors(g_bins, g_dirs, r_bins.astype(np.float64),
                                     r_dirs.astype(np.float64))
        assert eb <= ECHO_TOL and ed <= ECHO_TOL, (label, eb, ed)
        assert float(np.abs(g_bins).sum()) > 0.0, label
    # the third call repeats the first: same rays
    assert np.array_equal(got[0][0], got[2][0])
    scene.close()


def test_noalloc_graph_replays_equal_the_eager_call(ctx):
    """The allocation-free launch captured in a HIP graph (the callers it is
    for) and replayed three times, with eager work between the replays:
    every replay equals the eager call ray for ray.  The jobs hand few rays
    to the 16-lane kernel, so its per-SIMD admission turns waves away and
    every replay depends on the counters being cleared again."""
    ctx.set_launch_config()
    ctx.set_handoff()
    ctx.set_handoff_lanes()
    ctx.set_route_rays()
    dev = torch.device("cuda", 0)
    mb, s1, l1, _ = scenes.shoebox(seed=1)
    box = H.Scene.from_builder(ctx, mb)
    mb, src, lis = scenes.c2_workload()
    city = H.Scene.from_builder(ctx, mb)
    jobs = [("shoebox 1 x 4096 rays", box,
             H.trace_params(4096, 3, max_bounces=64, seed=0xC1, bin_dt=0.001, max_time=2.0,
                            floor_rel=1e-6), np.asarray([s1], np.float32), np.asarray([l1], np.float32)),
            ("city block 16 x 512 rays x 8 listeners", city,
             H.trace_params(512, 3, seed=7, bin_dt=0.001, max_time=2.0), src, lis)]
    stream = torch.cuda.Stream(device=dev)
    for label, scene, p, s, l in jobs:
        d_src = torch.from_numpy(np.ascontiguousarray(s, np.float32)).to(dev)
        d_lis = torch.from_numpy(np.ascontiguousarray(l, np.float32)).to(dev)
        e_rec, e_bins, e_dirs = _host(_launch(H.trace_world_async_noalloc, scene, p, d_src, d_lis, (),
                                              stream))
        bins, dirs, recs = _buffers(p, d_src.shape[0], d_lis.shape[0])
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=stream):
            H.trace_world_async_noalloc(scene, p, d_src.data_ptr(), d_src.shape[0], d_lis.data_ptr(),
                                        d_lis.shape[0], bins.data_ptr(), dirs.data_ptr(),
                                        stream=stream.cuda_stream, d_records=recs.data_ptr())
        for rep in range(3):
            # eager work between the replays: the outputs poisoned again
            bins.fill_(float("nan"))
            dirs.fill_(float("nan"))
            recs.zero_()
            torch.cuda.synchronize()
            graph.replay()
            g_rec, g_bins, g_dirs = _host((bins, dirs, recs))
            print(f"case no-alloc graph replay {rep}: {label}", flush=True)
            diff = np.nonzero(np.any(g_rec != e_rec, axis=1))[0]
            assert diff.size == 0, f"{label}, replay {rep}: {diff.size} rays differ, first {diff[:5]}"
            assert not np.isnan(g_bins).any() and not np.isnan(g_dirs).any(), (label, rep)
            eb, ed = echogram_rel_errors(g_bins, g_dirs, e_bins.astype(np.float64),
                                         e_dirs.astype(np.float64))
            assert eb <= ECHO_TOL and ed <= ECHO_TOL, (label, rep, eb, ed)
        del graph
    box.close()
    city.close()


def test_noalloc_refuses_beyond_its_limits(ctx):
    dev = torch.device("cuda", 0)
    mb, src, lis = scenes.c2_workload()
    scene = H.Scene.from_builder(ctx, mb)
    d_src = torch.from_numpy(np.ascontiguousarray(src[:4], np.float32)).to(dev)
    d_lis = torch.from_numpy(np.ascontiguousarray(lis[:1], np.float32)).to(dev)
    p = H.trace_params(65537, 3, bin_dt=0.001, max_time=0.5)   # 4 x 65537 > 262144 rays
    bins, dirs, recs = _buffers(H.trace_params(1, 3, bin_dt=0.001, max_time=0.5), 4, 1)
    stream = torch.cuda.Stream(device=dev)
    with pytest.raises(H.OmgError) as e:
        H.trace_world_async_noalloc(scene, p, d_src.data_ptr(), 4, d_lis.data_ptr(), 1,
                                    bins.data_ptr(), dirs.data_ptr(), stream=stream.cuda_stream)
    assert e.value.code == H._lib.ERR_CAP
    assert "262144" in str(e.value)
    scene.close()
