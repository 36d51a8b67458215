"""Host-side preparation of the device operator (structure only, no arithmetic
of the solve): directed faces, boundary lists, contact data and the
per-direction sweep schedules from the native scheduler.

Follows the index semantics of openbte/bte.py:47-158: directed internal
faces (i -> j with k = nA/V_i, j -> i with k = -nA/V_j), diffuse walls of
``side_physical_regions[bcs.diffuse]``, isothermal walls of ``bcs.mixed``,
periodic jumps on ``periodic_sides`` and the kappa contact.
"""
from __future__ import annotations

import ctypes as C
from typing import NamedTuple, Optional

import numpy as np

from . import _lib
from .cache import SCHEDULES, content_key


class HostOperator(NamedTuple):
    n: int
    dim: int
    n_mfp: int
    n_dirs: int
    # directed internal faces
    i_mat: np.ndarray
    j_mat: np.ndarray
    k_face: np.ndarray
    jump_face: np.ndarray
    cut_candidate: np.ndarray
    cf_ptr: np.ndarray
    cf_face: np.ndarray
    # walls per cell (diffuse and mixed)
    wall_ptr: np.ndarray
    wall_side: np.ndarray
    wall_db: np.ndarray
    wall_temp: np.ndarray
    # diffuse wall sides
    i_diffuse: np.ndarray
    wall_dbs: np.ndarray
    wall_rinv: np.ndarray
    # kappa contact
    contact_kind: int
    ce1: np.ndarray
    ce2: np.ndarray
    c_na: np.ndarray
    contact_jump: float
    normalization: float
    # material
    dirs: np.ndarray
    mfps: np.ndarray
    t_coeff: np.ndarray
    sigma: np.ndarray
    heat: np.ndarray
    nano: float
    # partially transmitting interfaces (None: every face transmits fully)
    t_face: Optional[np.ndarray] = None   # (F,) transmission of each directed face
    c_t: Optional[np.ndarray] = None      # (n_contact,) transmission of each contact side


class Schedule(NamedTuple):
    order: np.ndarray
    pos: np.ndarray
    lvl_off: np.ndarray
    lvl_ptr: np.ndarray
    in_ptr: np.ndarray
    in_off: np.ndarray
    in_src: np.ndarray
    in_face: np.ndarray
    cut_off: np.ndarray
    cut_exit: np.ndarray
    cut_face: np.ndarray
    n_levels: np.ndarray
    # fixed-width inflow slots [L][kin][n] (see fixed_width_inflow), exported natively
    kin: int = 0
    nb_src: Optional[np.ndarray] = None
    nb_face: Optional[np.ndarray] = None


def n_directions(mat) -> int:
    g = mat.grid
    return int(g[1] * g[2]) if len(g) == 3 else int(g[1])


def build_host_operator(geo, mat, bcs, effective_thermal_conductivity=None, heat_source=None) -> HostOperator:
    dim = int(geo.dim)
    n = int(geo.n_elems)
    M = int(mat.grid[0])
    L = n_directions(mat)
    sem = np.asarray(geo.side_elem_map)
    na = np.asarray(geo.normal_areas, dtype=np.float64)
    vol = np.asarray(geo.elem_volumes, dtype=np.float64)
    inn = np.asarray(geo.internal, dtype=np.int64)
    i, j = sem[inn, 0], sem[inn, 1]
    i_mat = np.concatenate([i, j]).astype(np.int32)
    j_mat = np.concatenate([j, i]).astype(np.int32)
    k_face = np.ascontiguousarray(np.concatenate([na[inn] / vol[i][:, None], -na[inn] / vol[j][:, None]]))
    F = len(i_mat)
    nint = len(inn)
    # periodic faces: cut candidates and jump with the sign of openbte/bte.py:106,110
    side_jump = np.zeros(len(geo.sides))
    is_per = np.zeros(len(geo.sides), bool)
    for region, sides in geo.periodic_sides.items():
        sides = np.asarray(sides, dtype=np.int64)
        is_per[sides] = True
        side_jump[sides] = float(bcs.periodic.get(region, 0.0)) if region in bcs.periodic else 0.0
    jump_face = np.concatenate([-side_jump[inn], side_jump[inn]])
    # the reference applies a jump only for regions listed in bcs.periodic
    cut_candidate = np.concatenate([is_per[inn], is_per[inn]]).astype(np.uint8)
    order = np.argsort(i_mat, kind='stable')
    cf_ptr = np.zeros(n + 1, np.int32)
    np.add.at(cf_ptr, i_mat + 1, 1)
    cf_ptr = np.cumsum(cf_ptr).astype(np.int32)
    cf_face = order.astype(np.int32)

    # walls: diffuse (index into the diffuse list) then isothermal (-1)
    s = np.asarray(mat.vmfp, dtype=np.float64)
    wc, wside, wdb, wtemp = [], [], [], []
    if len(bcs.diffuse) > 0:
        s_diffuse = np.asarray(geo.side_physical_regions[bcs.diffuse], dtype=np.int64)
        i_diffuse = sem[s_diffuse, 0]
        dbs = na[s_diffuse] / vol[i_diffuse][:, None]
        # tot_s = sum over the modes of min(sigma . dbs_s, 0), a chunk of
        # sides at a time (the full modes x sides array runs to GBs on 3D
        # meshes; the values per side are the same)
        sig = np.ascontiguousarray(np.asarray(mat.sigma)[:, :, :dim])
        tot = np.empty(len(s_diffuse))
        for c0 in range(0, len(s_diffuse), 512):
            Gb = np.einsum('mqj,sj->mqs', sig, dbs[c0:c0 + 512])
            tot[c0:c0 + 512] = np.minimum(Gb, 0.0, out=Gb).sum(axis=(0, 1))
        rinv = np.divide(1.0, tot, out=np.zeros_like(tot), where=tot != 0)
        wc.append(i_diffuse)
        wside.append(np.arange(len(s_diffuse)))
        wdb.append(dbs)
        wtemp.append(np.zeros(len(s_diffuse)))
    else:
        i_diffuse = np.zeros(0, np.int64)
        dbs = np.zeros((0, dim))
        rinv = np.zeros(0)
    for key, contact in bcs.mixed.items():
        sides = np.asarray(geo.side_physical_regions[key], dtype=np.int64)
        ii = sem[sides, 0]
        wc.append(ii)
        wside.append(-np.ones(len(sides), np.int64))
        wdb.append(na[sides] / vol[ii][:, None])
        wtemp.append(np.full(len(sides), float(contact['value'])))
    if wc:
        wc = np.concatenate(wc)
        wo = np.argsort(wc, kind='stable')
        wall_side = np.concatenate(wside)[wo].astype(np.int32)
        wall_db = np.ascontiguousarray(np.concatenate(wdb)[wo])
        wall_temp = np.concatenate(wtemp)[wo]
        wall_ptr = np.zeros(n + 1, np.int64)
        np.add.at(wall_ptr, wc + 1, 1)
        wall_ptr = np.cumsum(wall_ptr).astype(np.int32)
    else:
        wall_side = np.zeros(0, np.int32)
        wall_db = np.zeros((0, dim))
        wall_temp = np.zeros(0)
        wall_ptr = np.zeros(n + 1, np.int32)

    contact_kind, ce1, ce2, c_na, alpha, normalization = 0, np.zeros(0, np.int64), np.zeros(0, np.int64), \
        np.zeros((0, dim)), 0.0, 0.0
    if effective_thermal_conductivity is not None:
        if bcs.is_nanowire:
            raise NotImplementedError('nanowire kappa contact (openbte/bte.py:138-143) is not supported')
        contact = effective_thermal_conductivity.contact
        normalization = float(effective_thermal_conductivity.normalization)
        if bcs.get_bc_by_name(contact) == 'Mixed':
            sides = np.asarray(geo.side_physical_regions[contact], dtype=np.int64)
            contact_kind, alpha = 2, float(bcs.mixed[contact]['value'])
            ce1 = sem[sides, 0]
        else:
            sides = np.asarray(geo.periodic_sides[contact], dtype=np.int64)
            contact_kind, alpha = 1, float(bcs.periodic[contact])
            ce1, ce2 = sem[sides, 0], sem[sides, 1]
        c_na = na[sides]

    if heat_source is None:
        hs = np.zeros(n)
    elif hasattr(heat_source, 'variables'):
        hs = np.asarray(heat_source.variables['Heat_Source']['data'], dtype=np.float64)
    else:
        hs = np.asarray(heat_source, dtype=np.float64)
    heat = hs * float(mat.heat_source_coeff)

    return HostOperator(
        n=n, dim=dim, n_mfp=M, n_dirs=L,
        i_mat=i_mat, j_mat=j_mat, k_face=k_face, jump_face=jump_face, cut_candidate=cut_candidate,
        cf_ptr=cf_ptr, cf_face=cf_face,
        wall_ptr=wall_ptr, wall_side=wall_side, wall_db=wall_db, wall_temp=wall_temp,
        i_diffuse=np.asarray(i_diffuse, np.int64), wall_dbs=np.ascontiguousarray(dbs), wall_rinv=rinv,
        contact_kind=contact_kind, ce1=np.asarray(ce1, np.int32), ce2=np.asarray(ce2, np.int32),
        c_na=np.ascontiguousarray(c_na), contact_jump=alpha, normalization=normalization,
        dirs=np.ascontiguousarray(s[:, :dim]), mfps=np.asarray(mat.mfps, np.float64),
        t_coeff=np.ascontiguousarray(np.asarray(mat.t_coeff, np.float64)),
        sigma=np.ascontiguousarray(np.asarray(mat.sigma, np.float64)[:, :, :dim]),
        heat=heat, nano=1.0 if bcs.is_nanowire else 0.0)


# Cells per level of the default schedule: with at most 192 cells a level
# fits one pass of a 256-thread windowed sweep workgroup at mc = 8 (3 cells
# per thread) and two adjacent levels fit a 512-position LDS window.
DEFAULT_MAX_WIDTH = 192


def separable_velocities(sigma: np.ndarray, rtol: float = 1e-13):
    """``(sig_a [M][L], sig_dir [L][dim])`` with sigma_ml = sig_a_ml sig_dir_l
    (sig_dir_l a unit vector) when every direction's group velocities are
    parallel to within ``rtol`` of max |sigma|, else ``(None, None)``.  Gray
    models give such velocities (openbte/material.py:192-305: sigma = const
    s_l); the RTA models' velocities, interpolated from the bulk modes onto the
    angle grid (openbte/material.py:7-189), generally do not.  With them the
    flux of a position needs one mode sum instead of dim."""
    sigma = np.asarray(sigma, np.float64)
    M, L, dim = sigma.shape
    scale = np.abs(sigma).max() if sigma.size else 0.0
    if scale == 0.0:
        return None, None
    nrm = np.linalg.norm(sigma, axis=2)
    top = nrm.argmax(axis=0)
    ref = sigma[top, np.arange(L)]
    rn = nrm[top, np.arange(L)][:, None]
    d = np.divide(ref, rn, out=np.zeros_like(ref), where=rn > 0)
    a = np.einsum('mld,ld->ml', sigma, d)
    if np.abs(sigma - a[:, :, None] * d[None]).max() > rtol * scale:
        return None, None
    return np.ascontiguousarray(a), np.ascontiguousarray(d)


def batch_max_width(mc: int, block: int) -> int:
    """Cells per level of the schedules batched and served problems use: one
    pass of the persistent batch kernel, 2 cells per thread and mc / 2 lanes
    per cell (128 cells at mc = 8, block = 256).  Wider levels would take a
    second pass that reloads its records without prefetch.  The solution does
    not depend on the layering (cut faces are numbered by face id), so these
    problems give the same bits as with the default schedule."""
    return max(16, 4 * int(block) // int(mc))


def build_schedule(op: HostOperator, dirs: Optional[np.ndarray] = None, cut_candidate: Optional[np.ndarray] = None,
                   max_width: int = DEFAULT_MAX_WIDTH) -> Schedule:
    """Run the native level scheduler (libopenbte_amd obte_sched_*).
    ``max_width`` caps the cells of one level (0: longest-path levels)."""
    dirs = op.dirs if dirs is None else np.ascontiguousarray(dirs, dtype=np.float64)
    cut = op.cut_candidate if cut_candidate is None else np.ascontiguousarray(cut_candidate, dtype=np.uint8)
    # the schedule depends on the faces, the cut set and the directions only:
    # solves that reuse a mesh reuse it (openbte_amd.cache)
    key = content_key('schedule', op.n, op.dim, np.asarray(op.i_mat), np.asarray(op.j_mat),
                      np.asarray(op.k_face), np.asarray(cut), np.asarray(dirs), int(max_width))
    return SCHEDULES.get_or_build(key, lambda: _build_schedule(op, dirs, cut, max_width))


def _build_schedule(op: HostOperator, dirs: np.ndarray, cut: np.ndarray, max_width: int) -> Schedule:
    L = len(dirs)
    n = op.n
    i_mat = np.ascontiguousarray(op.i_mat, np.int32)
    j_mat = np.ascontiguousarray(op.j_mat, np.int32)
    k = np.ascontiguousarray(op.k_face, np.float64)
    h = C.c_void_p()
    _lib.call('obte_sched_create', n, len(i_mat), op.dim, _lib.ptr(i_mat), _lib.ptr(j_mat), _lib.ptr(k),
              _lib.ptr(cut), L, _lib.ptr(dirs), int(max_width), C.byref(h))
    try:
        sizes = np.zeros(3 * L, np.int64)
        _lib.call('obte_sched_sizes', h, _lib.ptr(sizes))
        sizes = sizes.reshape(L, 3)
        nlev, nent, ncut = sizes[:, 0], sizes[:, 1], sizes[:, 2]
        order = np.zeros(L * n, np.int32)
        pos = np.zeros(L * n, np.int32)
        lvl_ptr = np.zeros(int((nlev + 1).sum()), np.int32)
        in_ptr = np.zeros(L * (n + 1), np.int32)
        in_src = np.zeros(max(1, int(nent.sum())), np.int32)
        in_face = np.zeros(max(1, int(nent.sum())), np.int32)
        cut_exit = np.zeros(max(1, int(ncut.sum())), np.int32)
        cut_face = np.zeros(max(1, int(ncut.sum())), np.int32)
        _lib.call('obte_sched_export', h, _lib.ptr(order), _lib.ptr(pos), _lib.ptr(lvl_ptr), _lib.ptr(in_ptr),
                  _lib.ptr(in_src), _lib.ptr(in_face), _lib.ptr(cut_exit), _lib.ptr(cut_face))
        kin = max(2, int(np.diff(in_ptr.reshape(L, n + 1), axis=1).max(initial=0)))
        nb_src = np.empty((L, kin, n), np.int32)
        nb_face = np.empty((L, kin, n), np.int32)
        _lib.call('obte_sched_export_slots', h, kin, _lib.ptr(nb_src), _lib.ptr(nb_face))
    finally:
        _lib.call('obte_sched_destroy', h)
    lvl_off = np.concatenate([[0], np.cumsum(nlev + 1)]).astype(np.int32)
    in_off = np.concatenate([[0], np.cumsum(nent)]).astype(np.int64)
    cut_off = np.concatenate([[0], np.cumsum(ncut)]).astype(np.int32)
    return Schedule(order=order, pos=pos, lvl_off=lvl_off, lvl_ptr=lvl_ptr, in_ptr=in_ptr, in_off=in_off,
                    in_src=in_src, in_face=in_face, cut_off=cut_off, cut_exit=cut_exit, cut_face=cut_face,
                    n_levels=nlev.astype(np.int64), kin=kin, nb_src=nb_src, nb_face=nb_face)


def level_pair_max(sched: Schedule) -> int:
    """Largest |level k| + |level k+1| over all directions (a one-level
    direction counts its single level)."""
    best = 0
    for l in range(len(sched.n_levels)):
        lp = sched.lvl_ptr[sched.lvl_off[l]:sched.lvl_off[l] + int(sched.n_levels[l]) + 1]
        w = np.diff(lp)
        if len(w) == 0:
            continue
        best = max(best, int((w[:-1] + w[1:]).max()) if len(w) > 1 else int(w[0]))
    return best


def max_inflow(sched: Schedule) -> int:
    """Largest number of inflow faces of one cell over all directions."""
    return int(np.diff(sched.in_ptr.reshape(len(sched.n_levels), -1), axis=1).max(initial=0))


def fixed_width_inflow(sched: Schedule, n: int, kin: int):
    """Per-direction fixed-width inflow slots [L][kin][n] in position order.

    Slot k of position p holds the k-th inflow entry of p: its upstream row
    (a position, or n + cut index for a cut face) and its directed face.
    Unused slots point at p itself with face -1 (their coefficient is 0).
    Every row lies in [0, n + cuts of the direction)."""
    if sched.nb_src is not None and sched.kin == kin:
        return sched.nb_src, sched.nb_face
    L = len(sched.n_levels)
    nb_src = np.empty((L, kin, n), np.int32)
    nb_face = np.full((L, kin, n), -1, np.int32)
    nb_src[:] = np.arange(n, dtype=np.int32)[None, None, :]
    in_ptr = sched.in_ptr.reshape(L, n + 1).astype(np.int64)
    for l in range(L):
        e0, e1 = int(sched.in_off[l]), int(sched.in_off[l + 1])
        ip = in_ptr[l]
        cnt = np.diff(ip)
        if cnt.size and cnt.max() > kin:
            raise ValueError(f'direction {l}: {cnt.max()} inflow faces exceed {kin} slots')
        rows = np.repeat(np.arange(n), cnt)
        k = np.arange(e1 - e0) - ip[rows]
        src = sched.in_src[e0:e1].astype(np.int64)
        src = np.where(src >= 0, src, n + (~src))
        nb_src[l, k, rows] = src
        nb_face[l, k, rows] = sched.in_face[e0:e1]
    return nb_src, nb_face
