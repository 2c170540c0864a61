"""Semi-dense keyframe seeds (the sparse map's pre-fill) on the GPU against
the oracle's restatement of src/seed_init.cu and src/lsd_semidense.cu:
gradient, the initial keyframe, and frame-by-frame updates with keyframe
switches (propagation), seeds, pre-fill and (inverse depth, variance) prior
compared bit for bit."""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle as O
from open_quadtree_mapping_amd._lib import check, lib
from open_quadtree_mapping_amd.scene import camera_for, forward_lateral_trajectory, lateral_trajectory, make_scene, \
    render

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def p(t):
    return C.c_void_p(t.data_ptr())


def stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def gpu_gradient(img_t):
    H, W = img_t.shape
    g = torch.empty((H, W, 2), dtype=torch.float32, device="cuda")
    check(lib().qm_sd_gradient(p(img_t), W, H, W, p(g), stream()))
    return g


@pytest.mark.parametrize("motion,W,H,seed", [("lateral", 376, 240, 31), ("forward", 376, 240, 32),
                                             ("forward", 752, 480, 33)])
def test_semidense_sequence_matches_oracle(motion, W, H, seed):
    cam = camera_for(W, H)
    n = 9
    traj = lateral_trajectory(n, step=0.04, seed=seed) if motion == "lateral" else \
        forward_lateral_trajectory(n, seed=seed)
    sc = make_scene(seed)
    imgs = [render(sc, cam, T)[0] for T in traj]
    camv = np.array([cam.fx, cam.fy, cam.cx, cam.cy], np.float32)
    # frame 0 is the keyframe
    key_t = dev(imgs[0])
    kg = gpu_gradient(key_t)
    kg_ref = O.sd_gradient(imgs[0])
    assert np.array_equal(kg.cpu().numpy(), kg_ref)
    seeds = torch.empty((H, W, 8), dtype=torch.float32, device="cuda")
    check(lib().qm_sd_init(p(kg), W, H, p(seeds), stream()))
    ref_seeds = O.sd_init(kg_ref)
    assert np.array_equal(seeds.cpu().numpy(), ref_seeds)
    key_T, key_img, key_grad_ref = traj[0], imgs[0], kg_ref
    prefill = torch.empty((H, W), dtype=torch.float32, device="cuda")
    prior = torch.empty((H, W, 2), dtype=torch.float32, device="cuda")
    switched = 0
    for k in range(1, n):
        print(f"semidense {motion} {W}x{H} frame {k}", flush=True)
        inc_t = dev(imgs[k])
        ig = gpu_gradient(inc_t)
        ig_ref = O.sd_gradient(imgs[k])
        sw = O.need_switch_keyframe(key_T, traj[k]) or k in (4, 7)  # also force switches
        Tk = np.ascontiguousarray(key_T, np.float32).reshape(12)
        Ti = np.ascontiguousarray(traj[k], np.float32).reshape(12)
        check(lib().qm_sd_step(p(seeds), p(key_t), p(kg), p(inc_t), p(ig), W, H, W,
                               camv.ctypes.data_as(C.c_void_p), Tk.ctypes.data_as(C.c_void_p),
                               Ti.ctypes.data_as(C.c_void_p), int(sw), p(prefill), p(prior), stream()))
        want_prefill, want_prior = O.sd_step(ref_seeds, key_img, key_grad_ref, imgs[k], ig_ref, cam, key_T, traj[k],
                                             sw, want_prior=True)
        got_seeds = seeds.cpu().numpy()
        bad = np.argwhere(~((got_seeds == ref_seeds) | (np.isnan(got_seeds) & np.isnan(ref_seeds))).all(axis=2))
        assert bad.size == 0, f"frame {k}: {len(bad)} seeds differ, first at {bad[:3].tolist()}"
        assert np.array_equal(prefill.cpu().numpy(), want_prefill), f"frame {k}: pre-fill differs"
        assert np.array_equal(prior.cpu().numpy(), want_prior), f"frame {k}: (inverse depth, variance) prior differs"
        if sw:
            switched += 1
            key_t, kg, key_T, key_img, key_grad_ref = inc_t, ig, traj[k], imgs[k], ig_ref
    assert switched >= 2
    assert (prefill.cpu().numpy() > 0).sum() > 100
    assert (prior.cpu().numpy()[..., 0] > 0).sum() > 100
