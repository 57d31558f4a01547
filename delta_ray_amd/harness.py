"""Inverse rendering with the GPU renderer: Adam on scene parameters against a
target image, as the reference's optimisation scripts do
(test/test_single_triangle.py:62-77, test_single_triangle_camera.py,
test_single_triangle_fisheye.py, test_shadow.py:66-84).

A task builds the scene from a dict of parameter tensors; ``optimize`` renders
the target from the task's target parameters (``target_spp`` samples, seed 0),
then runs Adam on a copy of the start parameters: each iteration renders with
``spp`` samples and a new seed through ``RenderFunction`` and takes the L2
loss against the target (the reference's ``(img - target).pow(2).sum()``).
Images are written with ``image.imwrite`` when ``out_dir`` is given (target.exr,
init.png, iter_<t>.png, final.png).  Parameters live on the GPU, so every
iteration refits the kept scene on the device.

    python -m delta_ray_amd.harness single_triangle --iters 200 --out results/single_triangle
"""
import argparse
import json
import os
import time

import numpy as np
import torch

from . import image
from . import scene as sc
from .render_pytorch import RenderFunction


def _t(x, dev):
    return torch.tensor(np.asarray(x, np.float32), device=dev)


class Task:
    """name; params (start values), target (target values): dicts of arrays;
    build(p) -> (camera, materials, shapes, lights) from tensors p; lr;
    resolution; max_bounces"""

    def __init__(self, name, params, target, build, lr, resolution=(256, 256), max_bounces=1):
        self.name, self.params, self.target, self.build = name, params, target, build
        self.lr, self.resolution, self.max_bounces = lr, tuple(resolution), max_bounces


def _camera(pos, look, res, fisheye=False):
    """the camera of parameter tensors (its matrices are built on the host,
    as the reference's camera.py does; autograd carries their gradients back
    to the tensors wherever they live)"""
    return sc.Camera(pos.cpu(), look.cpu(), torch.tensor([0.0, 1.0, 0.0]), None, 45.0, 0.01, 10000.0, list(res),
                     fisheye=fisheye)


def single_triangle(res=(256, 256)):
    """test_single_triangle.py: the triangle's vertices"""
    def build(p):
        dev = p["vertices"].device
        grey = sc.Material(torch.tensor([0.5, 0.5, 0.5]))
        ind = torch.tensor([[0, 1, 2]], dtype=torch.int32, device=dev)
        lv = _t([[-1, -1, -7], [1, -1, -7], [-1, 1, -7], [1, 1, -7]], dev)
        li = torch.tensor([[0, 1, 2], [1, 3, 2]], dtype=torch.int32, device=dev)
        shapes = [sc.Shape(p["vertices"], ind, None, None, 0), sc.Shape(lv, li, None, None, 0)]
        cam = _camera(_t([0, 0, -5], dev), _t([0, 0, 0], dev), res)
        return cam, [grey], shapes, [sc.Light(1, torch.tensor([20.0, 20.0, 20.0]))]
    return Task("single_triangle",
                {"vertices": [[-2.0, 1.5, 0.3], [0.9, 1.2, -0.3], [-0.4, -1.4, 0.2]]},
                {"vertices": [[-1.7, 1.0, 0.0], [1.0, 1.0, 0.0], [-0.5, -1.0, 0.0]]}, build, 5e-2, res)


def single_triangle_camera(res=(256, 256), fisheye=False):
    """test_single_triangle_camera.py (and _fisheye.py): the camera's
    position and look-at point"""
    def build(p):
        dev = p["position"].device
        grey = sc.Material(torch.tensor([0.5, 0.5, 0.5]))
        v = _t([[-1.7, 1.0, 0.0], [1.0, 1.0, 0.0], [-0.5, -1.0, 0.0]], dev)
        ind = torch.tensor([[0, 1, 2]], dtype=torch.int32, device=dev)
        lz = -9.0 if fisheye else -7.0
        lv = _t([[-1, -1, lz], [1, -1, lz], [-1, 1, lz], [1, 1, lz]], dev)
        li = torch.tensor([[0, 1, 2], [1, 3, 2]], dtype=torch.int32, device=dev)
        shapes = [sc.Shape(v, ind, None, None, 0), sc.Shape(lv, li, None, None, 0)]
        cam = _camera(p["position"], p["look_at"], res, fisheye)
        inten = 30.0 if fisheye else 20.0
        return cam, [grey], shapes, [sc.Light(1, torch.tensor([inten] * 3))]
    if fisheye:
        return Task("single_triangle_fisheye", {"position": [0.2, -0.2, -1.2], "look_at": [0.0, 0.0, 0.0]},
                    {"position": [0.0, 0.0, -1.0], "look_at": [0.0, 0.0, 0.0]}, build, 5e-2, res)
    return Task("single_triangle_camera", {"position": [0.0, 0.0, -3.0], "look_at": [-0.5, -0.5, 0.0]},
                {"position": [0.0, 0.0, -5.0], "look_at": [0.0, 0.0, 0.0]}, build, 5e-2, res)


def shadow(res=(256, 256)):
    """test_shadow.py: the blocker's vertices, seen through their shadow"""
    def build(p):
        dev = p["vertices"].device
        grey = sc.Material(torch.tensor([0.5, 0.5, 0.5]))
        black = sc.Material(torch.tensor([0.0, 0.0, 0.0]))
        quad = torch.tensor([[0, 1, 2], [1, 3, 2]], dtype=torch.int32, device=dev)
        fv = _t([[-2.0, 0.0, -2.0], [-2.0, 0.0, 2.0], [2.0, 0.0, -2.0], [2.0, 0.0, 2.0]], dev)
        lv = _t([[-0.1, 5, -0.1], [-0.1, 5, 0.1], [0.1, 5, -0.1], [0.1, 5, 0.1]], dev)
        li = torch.tensor([[0, 2, 1], [1, 2, 3]], dtype=torch.int32, device=dev)
        shapes = [sc.Shape(fv, quad, None, None, 0), sc.Shape(p["vertices"], quad, None, None, 0),
                  sc.Shape(lv, li, None, None, 1)]
        cam = _camera(_t([0, 2, -5], dev), _t([0, 0, 0], dev), res)
        return cam, [grey, black], shapes, [sc.Light(2, torch.tensor([1000.0, 1000.0, 1000.0]))]
    return Task("shadow",
                {"vertices": [[-0.2, 3.5, -0.8], [-0.8, 3.0, 0.3], [0.4, 2.8, -0.8], [0.3, 3.2, 1.0]]},
                {"vertices": [[-0.5, 3.0, -0.5], [-0.5, 3.0, 0.5], [0.5, 3.0, -0.5], [0.5, 3.0, 0.5]]},
                build, 1e-2, res)


TASKS = {"single_triangle": single_triangle, "single_triangle_camera": single_triangle_camera,
         "single_triangle_fisheye": lambda res=(256, 256): single_triangle_camera(res, fisheye=True),
         "shadow": shadow}


def _render(task, p, spp, seed):
    cam, mats, shapes, lights = task.build(p)
    args = RenderFunction.serialize_scene(cam, mats, shapes, lights, list(task.resolution), spp, task.max_bounces)
    return RenderFunction.apply(seed, *args)


def optimize(task, iters=200, spp=4, target_spp=256, out_dir=None, device="cuda", log=None):
    """Adam on task.params against the image of task.target; returns the
    history (loss per iteration, parameter distance to the target, seconds)."""
    dev = torch.device(device)
    target_p = {k: _t(v, dev) for k, v in task.target.items()}
    with torch.no_grad():
        target = _render(task, target_p, target_spp, 0).detach().to(dev)
    p = {k: _t(v, dev).requires_grad_(True) for k, v in task.params.items()}
    if out_dir:
        os.makedirs(out_dir, exist_ok=True)
        image.imwrite(target.cpu().numpy(), os.path.join(out_dir, "target.exr"))
        image.imwrite(target.cpu().numpy(), os.path.join(out_dir, "target.png"))
        with torch.no_grad():
            image.imwrite(_render(task, p, target_spp, 1).cpu().numpy(), os.path.join(out_dir, "init.png"))
    opt = torch.optim.Adam(list(p.values()), lr=task.lr)

    def distance():
        return float(sum(torch.linalg.norm(p[k].detach() - target_p[k]) ** 2 for k in p) ** 0.5)
    hist = {"loss": [], "param_distance": [distance()]}
    t0 = time.perf_counter()
    for t in range(iters):
        opt.zero_grad()
        img = _render(task, p, spp, t + 1).to(dev)
        loss = (img - target).pow(2).sum()
        loss.backward()
        opt.step()
        hist["loss"].append(float(loss.detach()))
        hist["param_distance"].append(distance())
        if out_dir:
            image.imwrite(img.detach().cpu().numpy(), os.path.join(out_dir, "iter_%d.png" % t))
        if log:
            log("%s iter %d loss %.5g distance %.4g" % (task.name, t, hist["loss"][-1], hist["param_distance"][-1]))
    torch.cuda.synchronize()
    hist["seconds"] = time.perf_counter() - t0
    hist["params"] = {k: v.detach().cpu().numpy().tolist() for k, v in p.items()}
    if out_dir:
        with torch.no_grad():
            image.imwrite(_render(task, p, target_spp, iters + 1).cpu().numpy(), os.path.join(out_dir, "final.png"))
    return hist


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("task", choices=sorted(TASKS))
    ap.add_argument("--iters", type=int, default=200)
    ap.add_argument("--spp", type=int, default=4)
    ap.add_argument("--target-spp", type=int, default=256)
    ap.add_argument("--res", type=int, default=256)
    ap.add_argument("--out", default=None, help="folder for target / iteration images")
    a = ap.parse_args()
    task = TASKS[a.task]((a.res, a.res))
    h = optimize(task, a.iters, a.spp, a.target_spp, a.out, log=print)
    print(json.dumps({"task": a.task, "first_loss": h["loss"][0], "last_loss": h["loss"][-1],
                      "start_distance": h["param_distance"][0], "final_distance": h["param_distance"][-1],
                      "seconds": round(h["seconds"], 3), "params": h["params"]}))


if __name__ == "__main__":
    main()
