"""Pins of the oracle's ray marcher (src/raymarching.cuh restated in oracle/oracle_render.hpp),
the reference the GPU renderer is checked against bit for bit.

Closed forms a plausible mistake breaks: Beer-Lambert opacity of a constant slab along an axis
ray (front-to-back compositing, the 0.5-cell march from max(dNear, Z_NEAR) to the far face, the
exponent's dx factor); an empty domain stays black; the floor's colour from its illumination,
checkerboard and the hit point; the debug slice samples the plane z = gz/2 where trilinear
interpolation reproduces a linear density exactly (precise) or takes the containing cell
(coarse); the velocity and resolution channels; the sphere silhouette against the analytic
ray-sphere intersection."""
import ctypes as C
import math

import numpy as np
import pytest

import dcgrid_amd
from dcgrid_amd import scenes
from dcgrid_amd._abi import (RENDER_DENSITY, RENDER_FLUID, RENDER_RESOLUTION, RENDER_VELOCITY, RenderParams)


def params_for(n, dx=1.0):
    p = scenes.reference_params(n, kind="uniform")
    p.dx, p.rdx = dx, 1.0 / dx
    return p


def base_render(w=2, h=2):
    r = RenderParams()
    dcgrid_amd.library().dcg_default_render_params(None, C.byref(r))
    r.width, r.height = w, h
    r.solids, r.shadows = 0, 0
    r.eye_dir[0], r.eye_dir[1], r.eye_dir[2] = 0.0, 0.0, 1.0
    r.fov = 0.5
    return r


def uniform_sim(oracle_lib, n, density=None, velocity=None, dx=1.0, solids=0):
    p = params_for(n, dx)
    p.enable_additional_solids = solids
    s = oracle_lib.OracleSim("uniform", (n, n, n), p, precision=32)
    if density is not None:
        s.import_field("density", np.broadcast_to(np.asarray(density, np.float32), (n ** 3,)).copy())
    if velocity is not None:
        s.import_field("velocity", np.tile(np.asarray(velocity, np.float32), n ** 3))
    return s


def test_beer_lambert_slab(oracle_lib):
    n, rho, dx = 32, 0.004, 2.0
    s = uniform_sim(oracle_lib, n, density=rho, dx=dx)
    r = base_render()
    r.ambient = 1.0  # illumination = 1: colour = smoke colour
    r.eye_origin[0], r.eye_origin[1], r.eye_origin[2] = n / 2 * dx, n / 2 * dx, -20 * dx
    img = s.render(r)
    px = img[1, 1]  # the pixel whose ray is exactly +z: (x + 0 - cx) * s = 0 for x = 1, w = 2
    want = 1.0 - math.exp(-rho * dx * n)  # the ray crosses the slab z in [0, n)
    assert abs(px[3] - want) < 2e-6, (px[3], want)
    assert np.allclose(px[:3], 0.9, atol=1e-6)
    s.close()


def test_empty_domain_is_black(oracle_lib):
    s = uniform_sim(oracle_lib, 16, density=0.0)
    r = base_render(8, 6)
    r.eye_origin[0], r.eye_origin[1], r.eye_origin[2] = 8.0, 8.0, -30.0
    assert not s.render(r).any()
    s.close()


def test_floor_colour_and_checkerboard(oracle_lib):
    n = 16
    s = uniform_sim(oracle_lib, n, density=0.0)
    r = base_render(2, 2)
    r.solids = 1
    eye = np.array([5.0, 10.0, -7.0])
    r.eye_origin[0], r.eye_origin[1], r.eye_origin[2] = eye
    r.eye_dir[1], r.eye_dir[2] = -0.6, 0.8  # centre ray (0, -0.6, 0.8): hits y = 0 at t = 10/0.6
    sun = np.array([0.0, 0.8, 0.6])
    r.sun_dir[0], r.sun_dir[1], r.sun_dir[2] = sun
    img = s.render(r)
    t = 10.0 / 0.6
    hit = eye + t * np.array([0.0, -0.6, 0.8])
    checker = 1.0 + 0.1 * (abs(math.floor(hit[0] / n) + math.floor(hit[2] / n)) % 2)
    illum = (r.ambient + sun[1] * (1 - r.ambient)) * checker
    bg = 1.0 - math.exp(-1e-6 * 1.0 * t)
    # one composite onto transparent black: colour = source colour, alpha = 1 - bg
    assert np.allclose(img[1, 1, :3], illum * np.array(r.floor_color[:]), rtol=1e-5)
    assert abs(img[1, 1, 3] - (1 - bg)) < 1e-6
    s.close()


def linear_density(n, a=(0.01, -0.02, 0.005), c=0.5):
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    return (c + a[0] * (x + 0.5) + a[1] * (y + 0.5) + a[2] * (z + 0.5)).astype(np.float32).ravel()


@pytest.mark.parametrize("precise", [1, 0])
def test_density_slice(oracle_lib, precise):
    n = 32
    s = uniform_sim(oracle_lib, n, density=linear_density(n))
    r = base_render(16, 12)
    r.channel, r.precise = RENDER_DENSITY, precise
    r.eye_origin[0], r.eye_origin[1], r.eye_origin[2] = 15.3, 13.7, -40.0
    img = s.render(r)
    cx, cy = 8.0, 6.0
    sc = r.fov / max(cx, cy)
    checked = 0
    for y in range(r.height):
        for x in range(r.width):
            v = np.array([(x - cx) * sc, (y - cy) * sc, 1.0])
            d = v / np.linalg.norm(v)
            d = np.array([d[0], d[2] * 0.0 - d[1] * 1.0, d[1] * 0.0 + d[2] * 1.0])  # eye_dir (0, 0, 1)
            t = (n / 2 - (-40.0)) / d[2]
            p = np.array([15.3, 13.7, -40.0]) + t * d
            if not (0.6 <= p[0] <= n - 0.6 and 0.6 <= p[1] <= n - 0.6):
                continue
            if precise:
                want = 0.5 + 0.01 * p[0] - 0.02 * p[1] + 0.005 * (n / 2)
            else:
                i = np.floor(p).astype(int)
                want = 0.5 + 0.01 * (i[0] + 0.5) - 0.02 * (i[1] + 0.5) + 0.005 * (n / 2 + 0.5)
            assert abs(img[y, x, 0] - 100 * want) < 2e-3, (x, y, img[y, x, 0], 100 * want)
            assert img[y, x, 3] == 1.0 and img[y, x, 1] == 0.0
            checked += 1
    assert checked > 50
    s.close()


def test_velocity_and_resolution_channels(oracle_lib):
    n = 16
    u = (3.0, -4.0, 12.0)  # |u| = 13
    s = uniform_sim(oracle_lib, n, density=0.0, velocity=u)
    r = base_render(4, 4)
    r.eye_origin[0], r.eye_origin[1], r.eye_origin[2] = 6.2, 9.9, -10.0
    r.channel = RENDER_VELOCITY
    img = s.render(r)
    v = 13.0 * 0.1
    assert np.allclose(img[2, 2], [v - 2, v - 1, v, 1.0], atol=1e-5)
    r.channel, r.precise = RENDER_RESOLUTION, 0
    img = s.render(r)
    # the centre pixel's ray is +z through (6.2, 9.9): coarse cell (6, 9) -> (6/4 + 9/4) % 2 = 1
    assert np.array_equal(img[2, 2], np.array([1.0, 1.0, 1.0, 1.0], np.float32))
    s.close()


def test_sphere_silhouette(oracle_lib):
    n = 64
    s = uniform_sim(oracle_lib, n, density=0.0, dx=10000.0 / 64, solids=1)
    r = base_render(24, 24)
    r.solids = 1
    r.fov = 1.0
    eye = np.array([32.0, 28.8, -60.0])  # cells; the sphere is centred at (32, 28.8, 32)
    r.eye_origin[0], r.eye_origin[1], r.eye_origin[2] = eye * (10000.0 / 64)
    r.sun_dir[0], r.sun_dir[1], r.sun_dir[2] = 0.0, 0.0, -1.0  # towards the eye: lit front
    img = s.render(r)
    radius = 2000.0 / (10000.0 / 64)
    c = np.array([32.0, 28.8, 32.0])
    cx = cy = 12.0
    sc = r.fov / 12.0
    for y in range(24):
        for x in range(24):
            v = np.array([(x - cx) * sc, (y - cy) * sc, 1.0])
            d = v / np.linalg.norm(v)
            d = np.array([d[0], -d[1], d[2]])
            oc = eye - c
            b = np.dot(oc, d)
            disc = b * b - (np.dot(oc, oc) - radius ** 2)
            hit = disc > 0 and (-b - math.sqrt(disc)) > 0
            margin = abs(math.sqrt(max(np.dot(oc, oc) - b * b, 0)) - radius)
            if margin < 0.5:
                continue  # grazing rays: sphere tracing stops within 0.1 of the surface
            got = img[y, x, 3] > 0.5
            # downward rays that miss the sphere hit the floor (the sphere is above it)
            assert got == (hit or d[1] < 0), (x, y)
            if hit:
                assert np.allclose(img[y, x, :3], np.array(r.scene_color[:]) * img[y, x, :3].max() /
                                   max(r.scene_color[:]), rtol=1e-4)
    s.close()
