"""Heightfield WGAN with the differentiable renderer in the generator's path
(reference heightfield_gan.py:85-442).

A generator maps 128 normal variates to a 32 x 32 heightfield; the grid's
vertices are lifted by it, vertex normals follow (angle-weighted face
normals), and the mesh is rendered (32 x 32 pixels, 4 spp, direct light from
a small area light above) by RenderFunction from a camera on a circle of
radius 6 at height 3 looking at the centre.  A convolutional critic compares
the red channel of these images with renders of "real" heightfields, sums of
sines with random amplitude, frequencies and phase.  Training is WGAN with
gradient penalty (lambda 10, Adam 1e-4, betas 0.5 / 0.9) as in the reference;
the critic's gradient reaches the generator through the renderer's vertex
and normal adjoints.

    python -m delta_ray_amd.heightfield_gan --iters 1000 --out results/heightfield_gan
"""
import argparse
import math
import os
import time

import numpy as np
import torch
import torch.nn as nn

from . import image
from . import scene as sc
from .render_pytorch import RenderFunction

RES = 32      # heightfield grid and image size
NOISE = 128
LAMBDA = 10.0


def plane(n, extent=5.0):
    """an n x n grid over [-extent, extent]^2 at y = 0: vertices [n*n, 3],
    two triangles per cell [(n-1)^2 * 2, 3]"""
    t = np.linspace(-extent, extent, n, dtype=np.float32)
    X, Z = np.meshgrid(t, t, indexing="ij")
    v = np.stack([X, np.zeros_like(X), Z], -1).reshape(-1, 3)
    i, j = np.meshgrid(np.arange(n - 1), np.arange(n - 1), indexing="ij")
    a = (i * n + j).reshape(-1)
    b, c = a + 1, a + n
    d = c + 1
    ind = np.concatenate([np.stack([a, b, c], 1), np.stack([c, b, d], 1)]).astype(np.int32)
    return v, ind


def real_heights(n, rng, extent=5.0):
    """heights of a "real" sample: amplitude * sin(fx x + p) * fy * cos(z + p)
    with amplitude in [0.5, 0.7], fx, fy in [0.5, 1] and p in [0, pi] (the
    reference's generator of training heightfields, which uses the first phase
    in both factors)"""
    amp = 0.2 * rng.random() + 0.5
    fx, fy = 0.5 * rng.random() + 0.5, 0.5 * rng.random() + 0.5
    ph = math.pi * rng.random()
    t = np.linspace(-extent, extent, n, dtype=np.float32)
    X, Z = np.meshgrid(t, t, indexing="ij")
    return (amp * np.sin(fx * X + ph) * fy * np.cos(Z + ph)).astype(np.float32)


def vertex_normals(v, ind):
    """angle-weighted sums of face normals, normalised (differentiable)"""
    p = [v[ind[:, k]] for k in range(3)]
    fn = torch.cross(p[1] - p[0], p[2] - p[0], dim=1)
    fn = fn / torch.linalg.norm(fn, dim=1, keepdim=True).clamp_min(1e-12)
    n = torch.zeros_like(v)
    for k in range(3):
        a = p[(k + 1) % 3] - p[k]
        b = p[(k + 2) % 3] - p[k]
        a = a / torch.linalg.norm(a, dim=1, keepdim=True)
        b = b / torch.linalg.norm(b, dim=1, keepdim=True)
        # the corner angle as 2 asin(|b - a| / 2), stable for all angles
        ang = 2.0 * torch.asin((0.5 * torch.linalg.norm(b - a, dim=1)).clamp(max=1.0))
        n = n.index_add(0, ind[:, k], fn * ang[:, None])
    return n / torch.linalg.norm(n, dim=1, keepdim=True).clamp_min(1e-12)


class Generator(nn.Module):
    """noise [B, 128] -> heights [B, 32, 32] in [-1, 1]: a 4 x 4 x 256 seed
    upsampled twice by transposed convolutions; the heightfield is the mean
    of tanh outputs at 4, 8, 16 and 32 cells, each brought to 32 x 32"""

    def __init__(self, ngf=64):
        super().__init__()
        self.ngf = ngf
        self.seed = nn.Sequential(nn.Linear(NOISE, 4 * 4 * 4 * ngf), nn.ReLU(True))
        self.up8 = nn.Sequential(nn.ConvTranspose2d(4 * ngf, 2 * ngf, 2, stride=2, bias=False),
                                 nn.BatchNorm2d(2 * ngf), nn.ReLU(True))
        self.up16 = nn.Sequential(nn.ConvTranspose2d(2 * ngf, ngf, 2, stride=2, bias=False),
                                  nn.BatchNorm2d(ngf), nn.ReLU(True))
        self.out4 = nn.ConvTranspose2d(4 * ngf, 1, 1)
        self.out8 = nn.ConvTranspose2d(2 * ngf, 1, 1)
        self.out16 = nn.ConvTranspose2d(ngf, 1, 1)
        self.out32 = nn.ConvTranspose2d(ngf, 1, 2, stride=2)

    def forward(self, z):
        x4 = self.seed(z).view(-1, 4 * self.ngf, 4, 4)
        x8 = self.up8(x4)
        x16 = self.up16(x8)
        up = nn.functional.interpolate
        h = (torch.tanh(self.out32(x16)) +
             up(torch.tanh(self.out16(x16)), size=(RES, RES), mode="bilinear", align_corners=False) +
             up(torch.tanh(self.out8(x8)), size=(RES, RES), mode="bilinear", align_corners=False) +
             up(torch.tanh(self.out4(x4)), size=(RES, RES), mode="bilinear", align_corners=False)) / 4.0
        return h[:, 0]


class Critic(nn.Module):
    """image [B, 1, 32, 32] -> score [B, 1]"""

    def __init__(self, ndf=64):
        super().__init__()
        self.ndf = ndf
        self.conv = nn.Sequential(nn.Conv2d(1, ndf, 3, 2, 1), nn.LeakyReLU(0.01, inplace=True),
                                  nn.Conv2d(ndf, 2 * ndf, 3, 2, 1), nn.LeakyReLU(0.01, inplace=True),
                                  nn.Conv2d(2 * ndf, 4 * ndf, 3, 2, 1), nn.LeakyReLU(0.01, inplace=True))
        self.linear = nn.Linear(4 * 4 * 4 * ndf, 1)

    def forward(self, x):
        return self.linear(self.conv(x).view(x.shape[0], -1))


class Renderer:
    """renders heightfields [B, 32, 32] (device tensors) to red-channel
    images [B, 1, 32, 32] from given camera angles"""

    def __init__(self, dev, spp=4):
        self.dev, self.spp = dev, spp
        v, ind = plane(RES)
        self.base = torch.from_numpy(v).to(dev)
        self.ind = torch.from_numpy(ind).to(dev)
        self.ind_long = self.ind.long()
        lv = np.array([[-0.1, 10, -0.1], [-0.1, 10, 0.1], [0.1, 10, -0.1], [0.1, 10, 0.1]], np.float32)
        self.light = sc.Shape(torch.from_numpy(lv).to(dev),
                              torch.tensor([[0, 2, 1], [1, 2, 3]], dtype=torch.int32, device=dev), None, None, 1)
        self.materials = [sc.Material(torch.tensor([0.5, 0.5, 0.5])), sc.Material(torch.tensor([0.0, 0.0, 0.0]))]
        self.lights = [sc.Light(1, torch.tensor([4000.0, 4000.0, 4000.0]))]
        self.seed = 0

    def __call__(self, heights, angles):
        out = []
        for h, a in zip(heights, angles):
            lift = torch.stack([torch.zeros_like(h), h, torch.zeros_like(h)], -1).reshape(-1, 3)
            v = self.base + lift
            n = vertex_normals(v, self.ind_long)
            shapes = [sc.Shape(v, self.ind, None, n, 0), self.light]
            cam = sc.Camera(torch.tensor([6.0 * math.cos(a), 3.0, 6.0 * math.sin(a)]), torch.tensor([0.0, 0.0, 0.0]),
                            torch.tensor([0.0, 1.0, 0.0]), None, 45.0, 0.01, 10000.0, [RES, RES])
            args = RenderFunction.serialize_scene(cam, self.materials, shapes, self.lights, [RES, RES], self.spp, 1)
            self.seed += 1
            out.append(RenderFunction.apply(self.seed, *args)[:, :, 0])
        return torch.stack(out)[:, None]


def gradient_penalty(critic, real, fake, rng):
    alpha = torch.tensor(rng.random(), dtype=real.dtype, device=real.device)
    mix = (alpha * real + (1 - alpha) * fake).requires_grad_(True)
    score = critic(mix)
    g = torch.autograd.grad(score, mix, torch.ones_like(score), create_graph=True)[0]
    return LAMBDA * ((g.view(g.shape[0], -1).norm(2, dim=1) - 1) ** 2).mean()


def train(iters=1000, batch=4, out_dir=None, seed=0, log=print, dev="cuda"):
    """WGAN-GP steps; returns the per-iteration critic and generator costs"""
    dev = torch.device(dev)
    torch.manual_seed(seed)
    rng = np.random.default_rng(seed)
    G, D = Generator().to(dev), Critic().to(dev)
    for m in list(G.modules()) + list(D.modules()):
        if isinstance(m, (nn.Conv2d, nn.ConvTranspose2d)):
            nn.init.normal_(m.weight, 0.0, 0.02)
    optG = torch.optim.Adam(G.parameters(), lr=1e-4, betas=(0.5, 0.9))
    optD = torch.optim.Adam(D.parameters(), lr=1e-4, betas=(0.5, 0.9))
    render = Renderer(dev)
    hist = {"d_cost": [], "g_cost": [], "wasserstein": []}
    if out_dir:
        os.makedirs(out_dir, exist_ok=True)
    t0 = time.perf_counter()
    for it in range(iters):
        with torch.no_grad():
            real_h = torch.from_numpy(np.stack([real_heights(RES, rng) for _ in range(batch)])).to(dev)
            real = render(real_h, rng.uniform(0, 2 * math.pi, batch))
        # critic
        optD.zero_grad()
        noise = torch.randn(batch, NOISE, device=dev)
        angles = rng.uniform(0, 2 * math.pi, batch)
        with torch.no_grad():
            fake = render(G(noise), angles)
        d_real, d_fake = D(real).mean(), D(fake).mean()
        gp = gradient_penalty(D, real, fake, rng)
        d_cost = d_fake - d_real + gp
        d_cost.backward()
        optD.step()
        # generator, through the renderer
        optG.zero_grad()
        g_cost = -D(render(G(noise), angles)).mean()
        g_cost.backward()
        optG.step()
        hist["d_cost"].append(float(d_cost))
        hist["g_cost"].append(float(g_cost))
        hist["wasserstein"].append(float(d_real - d_fake))
        if log:
            log("iteration %d: D cost %.4g (W %.4g, GP %.4g), G cost %.4g" %
                (it, hist["d_cost"][-1], hist["wasserstein"][-1], float(gp), hist["g_cost"][-1]))
        if out_dir and (it % 100 == 0 or it == iters - 1):
            with torch.no_grad():
                G.eval()
                h = G(torch.randn(16, NOISE, device=dev, generator=torch.Generator(device=dev).manual_seed(1)))
                imgs = render(h, np.linspace(0, 2 * math.pi, 16, endpoint=False))
                G.train()
            tile = imgs[:, 0].reshape(4, 4, RES, RES).permute(0, 2, 1, 3).reshape(4 * RES, 4 * RES)
            image.imwrite(np.repeat(tile.cpu().numpy()[:, :, None], 3, axis=2),
                          os.path.join(out_dir, "generated_%06d.png" % it))
    hist["seconds"] = time.perf_counter() - t0
    hist["generator"], hist["critic"] = G, D
    return hist


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--iters", type=int, default=1000)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    h = train(a.iters, a.batch, a.out)
    print("trained %d iterations in %.1f s" % (a.iters, h["seconds"]))


if __name__ == "__main__":
    main()
