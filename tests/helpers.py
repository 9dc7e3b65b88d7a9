"""Shared helpers of the test suite: oracle/GPU result comparison and reference-test geometry."""
import math

import numpy as np


def dir_for_impact_parameter(D, b):
    """Camera on +z at distance D, ray in the xz-plane with impact parameter b
    (tests/geodesic_tests.cpp:26-30)."""
    s = b * math.sqrt(1.0 - 2.0 / D) / D
    a = math.asin(min(1.0, s))
    return np.array([math.sin(a), 0.0, -math.cos(a)])


def angle_between(a, b):
    """tests/geodesic_tests.cpp:32-34."""
    a = np.asarray(a, float) / np.linalg.norm(a)
    b = np.asarray(b, float) / np.linalg.norm(b)
    return math.acos(float(np.clip(np.dot(a, b), -1.0, 1.0)))


def oracle_to_dict(r):
    """Oracle LAWS_RAY_DTYPE array -> the field layout of RayBuffers.numpy()."""
    return dict(cls=r["cls"].astype(np.uint8), flags=r["flags"].astype(np.uint8),
                n_cross=np.minimum(r["n_cross"], 255).astype(np.uint8),
                steps=r["steps"].astype(np.int32), esc=r["esc"].astype(np.float32),
                disk_r=r["disk_r"].astype(np.float32), disk_phi=r["disk_phi"].astype(np.float32),
                disk_g=r["disk_g"].astype(np.float32), b=r["b"].astype(np.float32))


FIELDS = ("cls", "flags", "n_cross", "steps", "esc", "disk_r", "disk_phi", "disk_g", "b")


def bit_mismatches(gpu: dict, ref: dict, fields=FIELDS):
    """Per field, the number of rays whose value differs bit for bit."""
    out = {}
    for f in fields:
        a, b = np.asarray(gpu[f]), np.asarray(ref[f])
        if a.dtype == np.float32:
            a, b = a.view(np.uint32), b.view(np.uint32)
        d = a != b
        if d.ndim > 1:
            d = d.any(axis=1)
        out[f] = int(d.sum())
    return out


def escape_angle_error(esc_a, esc_b):
    """Angle (rad) between two sets of unit escape directions, computed in float64."""
    a = np.asarray(esc_a, np.float64)
    b = np.asarray(esc_b, np.float64)
    cr = np.linalg.norm(np.cross(a, b), axis=1)
    dt = np.sum(a * b, axis=1)
    return np.arctan2(cr, dt)


def oracle_params_for(cfg, root64=0, band_rel=0.02):
    """oracle.LawSParams with the settings of a workload dict (blackholerender_amd.workload)."""
    import oracle
    from blackholerender_amd.camera import PRESETS
    ss, dstep = PRESETS[cfg["max_steps"]]
    return oracle.lawS_params(max_steps=cfg["max_steps"], step_scale=ss, disk_step=dstep,
                              disk_inner=cfg["disk_inner"], disk_outer=cfg["disk_outer"],
                              disk_thickness=0.035, disk_mode=1, root64=root64, band_rel=band_rel)


def oracle_threads() -> int:
    """Worker threads for the CPU oracle: the CPUs this process may run on, at most 16."""
    import os
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def gather_rays(out, idx) -> dict:
    """Rays idx (a 1-D int64 index array) of device RayBuffers, gathered on the device and copied
    to the host in the field layout of RayBuffers.numpy()."""
    import torch
    i = torch.as_tensor(np.asarray(idx, np.int64), device=out.cls.device)
    d = out.disk[:, i].cpu().numpy()
    return dict(cls=out.cls[i].cpu().numpy(), flags=out.flags[i].cpu().numpy(),
                n_cross=out.n_cross[i].cpu().numpy(),
                steps=out.steps[i].cpu().numpy().view(np.uint16).astype(np.int32),
                esc=out.esc[:, i].cpu().numpy().T.copy(), disk_r=d[0:2].T.copy(),
                disk_phi=d[2:4].T.copy(), disk_g=d[4:6].T.copy(), b=out.b[i].cpu().numpy())
