"""Ab-initio RTA materials: database loading and interpolation onto the
(MFP bin, direction) grid of the solver.

``load_rta`` reads the reference's material databases (SQLite tables of
pickled numpy arrays, openbte/utils.py:147-151,260-279) WITHOUT executing the
pickles: a small interpreter walks the pickle opcodes (``pickletools``),
accepts only the shape of a pickled ``numpy.ndarray`` and rebuilds the array
from its raw bytes with ``numpy.frombuffer``.  Anything else in a record is
rejected.

``RTA2DSym`` and ``RTA3D`` restate openbte/material.py:105-189 and 7-102: the
modes are split bilinearly (2D: MFP x phi) or trilinearly (3D: MFP x theta x
phi) onto the grid by the HIP kernel ``obte_rta_scatter``; kappa, t_coeff,
the boundary conductance h and the heat-source ratio follow on the host.
"""
from __future__ import annotations

import os
import pickletools
import sqlite3
from typing import Optional

import numpy as np
import torch

from . import _lib
from .objects import Material, MaterialRTA

_FIELDS = ('heat_capacity', 'scattering_time', 'group_velocity', 'frequency', 'thermal_conductivity')


class _Global(tuple):
    pass


def decode_ndarray_pickle(blob: bytes) -> np.ndarray:
    """Rebuild the numpy array of a pickled ``ndarray`` without unpickling:
    the opcodes are interpreted as data (globals are never imported) and the
    result must have exactly the reconstruct/setstate shape of an ndarray."""
    stack, memo, marks = [], {}, []
    for op, arg, _ in pickletools.genops(blob):
        name = op.name
        if name in ('PROTO', 'FRAME'):
            continue
        if name in ('SHORT_BINUNICODE', 'BINUNICODE', 'UNICODE', 'BINUNICODE8', 'SHORT_BINSTRING', 'BINSTRING',
                    'BININT', 'BININT1', 'BININT2', 'LONG1', 'BINFLOAT', 'SHORT_BINBYTES', 'BINBYTES',
                    'BINBYTES8', 'BYTEARRAY8'):
            stack.append(bytes(arg) if isinstance(arg, (bytes, bytearray)) else arg)
        elif name in ('NEWTRUE', 'NEWFALSE', 'NONE'):
            stack.append({'NEWTRUE': True, 'NEWFALSE': False, 'NONE': None}[name])
        elif name == 'MEMOIZE':
            memo[len(memo)] = stack[-1]
        elif name in ('BINPUT', 'LONG_BINPUT', 'PUT'):
            memo[arg] = stack[-1]
        elif name in ('BINGET', 'LONG_BINGET', 'GET'):
            stack.append(memo[arg])
        elif name == 'STACK_GLOBAL':
            nm = stack.pop()
            mod = stack.pop()
            stack.append(_Global((mod, nm)))
        elif name == 'GLOBAL':
            mod, nm = arg.split(' ')
            stack.append(_Global((mod, nm)))
        elif name == 'MARK':
            marks.append(len(stack))
        elif name == 'TUPLE':
            k = marks.pop()
            t = tuple(stack[k:])
            del stack[k:]
            stack.append(t)
        elif name in ('TUPLE1', 'TUPLE2', 'TUPLE3'):
            k = {'TUPLE1': 1, 'TUPLE2': 2, 'TUPLE3': 3}[name]
            t = tuple(stack[-k:])
            del stack[-k:]
            stack.append(t)
        elif name == 'EMPTY_TUPLE':
            stack.append(())
        elif name == 'REDUCE':
            args = stack.pop()
            fn = stack.pop()
            stack.append(('call', fn, args))
        elif name == 'BUILD':
            state = stack.pop()
            obj = stack.pop()
            stack.append(('build', obj, state))
        elif name == 'STOP':
            break
        else:
            raise ValueError(f'pickle opcode {name} is not part of a pickled ndarray')
    if len(stack) != 1:
        raise ValueError('not a single pickled object')
    obj = stack[0]
    ok = (isinstance(obj, tuple) and obj[0] == 'build' and isinstance(obj[1], tuple) and obj[1][0] == 'call'
          and tuple(obj[1][1]) in (('numpy.core.multiarray', '_reconstruct'), ('numpy._core.multiarray', '_reconstruct')))
    if not ok:
        raise ValueError('the record is not a pickled numpy.ndarray')
    state = obj[2]
    _version, shape, dt, fortran, raw = state
    if not (isinstance(dt, tuple) and dt[0] == 'build' and dt[1][0] == 'call'
            and tuple(dt[1][1]) in (('numpy', 'dtype'),)):
        raise ValueError('unexpected dtype record')
    code = dt[1][2][0]
    order = dt[2][1]
    dtype = np.dtype(code).newbyteorder('<' if order == '<' else ('>' if order == '>' else '='))
    arr = np.frombuffer(raw, dtype=dtype).reshape(tuple(shape), order='F' if fortran else 'C')
    return arr.astype(dtype.newbyteorder('='), copy=True)


def load_rta(name: str, source: str = 'database', path: Optional[str] = None) -> MaterialRTA:
    """Load ``name`` (e.g. ``'Si_rta'``) like openbte/utils.py:147-151: from
    ``path`` (a directory) or the current directory (``source != 'database'``).
    Fields are taken in the table's row order, as the reference's positional
    ``MaterialRTA(*values)`` does; their keys are checked against that order."""
    base = path if path is not None else (os.getcwd() if source != 'database' else os.path.join(
        os.path.dirname(os.path.abspath(__file__)), 'materials'))
    fn = os.path.join(base, name + '.db')
    if not os.path.exists(fn):
        raise FileNotFoundError(fn)
    con = sqlite3.connect(f'file:{fn}?mode=ro', uri=True)
    try:
        rows = con.execute('SELECT key, value FROM unnamed ORDER BY rowid').fetchall()
    finally:
        con.close()
    keys = [k for k, _ in rows]
    if tuple(keys[:len(_FIELDS)]) != _FIELDS[:len(keys)]:
        raise ValueError(f'{fn}: unexpected fields {keys}')
    return MaterialRTA(*[decode_ndarray_pickle(v) for _, v in rows])


# -- interpolation --------------------------------------------------------------
def _modes(data: MaterialRTA, dim: int):
    """Per-mode W = C / tau, sigma = C v and vectorial MFP v tau; modes with
    |MFP| <= 1e-10 m dropped (openbte/material.py:32-49,125-148)."""
    C = np.asarray(data.heat_capacity, np.float64)
    tau = np.asarray(data.scattering_time, np.float64)
    v = np.asarray(data.group_velocity, np.float64)[:, :3]
    if dim == 2:
        v = np.concatenate([v[:, :2], np.zeros((len(v), 1))], axis=1)
    f = np.divide(1.0, tau, out=np.zeros_like(tau), where=tau != 0)
    W = C * f
    sigma = C[:, None] * v
    mfp = v * tau[:, None]
    r = np.linalg.norm(mfp, axis=1)
    keep = r > 1e-10
    return W[keep], sigma[keep], mfp[keep], r[keep]


def _scatter(W, sigma, mfp, dim, mfp_grid, phi, theta=None):
    dev = torch.device('cuda', torch.cuda.current_device())
    n_mfp, n_phi = len(mfp_grid), len(phi)
    n_theta = 1 if theta is None else len(theta)
    n_ang = n_phi * (n_theta if dim == 3 else 1)
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float64), device=dev)  # noqa: E731
    w_out = torch.zeros(n_mfp * n_ang, dtype=torch.float64, device=dev)
    s_out = torch.zeros(n_mfp * n_ang * dim, dtype=torch.float64, device=dev)
    Wd, Sd, Md, Gd, Pd = t(W), t(sigma), t(mfp), t(mfp_grid), t(phi)
    Td = t(theta if theta is not None else np.zeros(2))
    _lib.call('obte_rta_scatter', len(W), dim, Wd.data_ptr(), Sd.data_ptr(), Md.data_ptr(), n_mfp, Gd.data_ptr(),
              n_phi, Pd.data_ptr(), n_theta, Td.data_ptr(), w_out.data_ptr(), s_out.data_ptr(),
              torch.cuda.current_stream(dev).cuda_stream)
    return (w_out.view(n_mfp, n_ang).cpu().numpy(), s_out.view(n_mfp, n_ang, dim).cpu().numpy())


def _finish(Ws, Ss, mfp_grid, dirs, grid, h, coeff) -> Material:
    inv = np.divide(1.0, Ws, out=np.zeros_like(Ws), where=Ws != 0)
    kappa = np.einsum('mqi,mq,mqj->ij', Ss, inv, Ss)
    return Material(thermal_conductivity=kappa, sigma=Ss * 1e-9, vmfp=dirs, t_coeff=Ws / Ws.sum(),
                    mfps=mfp_grid * 1e9, grid=grid, h=h, heat_source_coeff=coeff)


def RTA2DSym(data: MaterialRTA, n_phi: int = 48, n_mfp: int = 50, mfp_min: Optional[float] = None,
             mfp_max: Optional[float] = None) -> Material:
    """In-plane material (openbte/material.py:105-189)."""
    W, sigma, mfp, r = _modes(data, 2)
    phi = np.linspace(0, 2.0 * np.pi, n_phi, endpoint=False)
    dirs = np.array([np.sin(phi), np.cos(phi)]).T
    mfp_max = np.max(r) * 1.1 if mfp_max is None else mfp_max
    mfp_min = np.min(r) * 0.9 if mfp_min is None else mfp_min
    # the 1.01 multiplies the exponent, as in openbte/material.py:153
    grid = np.logspace(np.log10(mfp_min) * 1.01, np.log10(mfp_max), n_mfp, endpoint=True)
    Ws, Ss = _scatter(W, sigma, mfp, 2, grid, phi)
    h = 2 / 3 * np.sum(W * mfp[:, 0].clip(min=0)) * 1e-9
    return _finish(Ws, Ss, grid, dirs, (n_mfp, n_phi), h, 1 / np.sum(W))


def RTA3D(data: MaterialRTA, n_phi: int = 24, n_mfp: int = 50, n_theta: int = 24) -> Material:
    """Full-sphere material (openbte/material.py:7-102)."""
    W, sigma, mfp, r = _modes(data, 3)
    dphi = 2.0 * np.pi / n_phi
    phi = np.linspace(dphi / 2.0, 2.0 * np.pi - dphi / 2.0, n_phi, endpoint=True)
    dth = np.pi / n_theta
    theta = np.linspace(dth / 2, np.pi - dth / 2, n_theta, endpoint=True)
    dirs = np.einsum('tj,pj->tpj', np.array([np.sin(theta), np.sin(theta), np.cos(theta)]).T,
                     np.array([np.sin(phi), np.cos(phi), np.ones(n_phi)]).T).reshape(-1, 3)
    grid = np.logspace(np.log10(np.min(r) * 0.9) * 1.01, np.log10(np.max(r) * 1.1), n_mfp, endpoint=True)
    Ws, Ss = _scatter(W, sigma, mfp, 3, grid, phi, theta)
    h = 2 / 3 * np.sum(W * mfp[:, 0].clip(min=0)) * 1e-18
    return _finish(Ws, Ss, grid, dirs, (n_mfp, n_theta, n_phi), h, 1 / np.sum(W))
