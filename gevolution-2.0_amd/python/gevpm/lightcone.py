"""Lightcone particle output on the device: the Gadget-2 branch of
writeLightcones (output.hpp:966-2160) for one rank, over the C ABI.

  Geometry        lightcone_geometry (metadata.hpp:205-212)
  vertices        findIntersectingLightcones (tools.hpp:619-862)
  IDLog           LightconeIDBacklog + IDprelog (lightcone_id_backlog.hpp,
                  output.hpp:1017-1185, 2166-2880)
  write_particles the per-cycle Gadget-2 lightcone files of one species
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import GevError, check, gev_gadget2_header, gev_lightcone_geometry, lib

MAX_INTERSECTS = 12           # metadata.hpp:34-36
MAX_VERTICES = 32             # vertices one output call accepts
IDCHECK_ZONE = 0.05           # LIGHTCONE_IDCHECK_ZONE, metadata.hpp:38-40
C_RHO_CRIT = 2.77459457e11    # metadata.hpp:98
GADGET_LENGTH_CONVERSION = 0.001  # metadata.hpp:111
GADGET_MASS_CONVERSION = 1.0e10   # metadata.hpp:114


@dataclass
class Geometry:
    """one lightcone: observer vertex (box units), redshift of the observer,
    direction, opening = cos(half-angle) (-1 = full sky), distance = (far,
    near) edges in box units"""
    vertex: tuple = (0.5, 0.5, 0.5)
    z: float = 0.0
    direction: tuple = (0.0, 0.0, 1.0)
    opening: float = -1.0
    distance: tuple = (1.0, 0.0)

    def c(self):
        g = gev_lightcone_geometry()
        for d in range(3):
            g.vertex[d] = self.vertex[d]
            g.direction[d] = self.direction[d]
        g.z = self.z
        g.opening = self.opening
        g.distance[0], g.distance[1] = self.distance
        return g


def box_domain(N):
    """writeLightcones' domain of a rank that holds the whole box
    (output.hpp:1175-1190): half a cell beyond each face, box units"""
    return np.array([-0.5 / N] * 3 + [(N + 0.5) / N] * 3)


def vertices(geo, outer, inner, domain, cap=MAX_VERTICES):
    """periodic copies of the observer whose shell [inner, outer] may meet the
    domain: (count, vertices[min(count, cap), 3])"""
    g = geo.c() if isinstance(geo, Geometry) else geo
    dom = np.ascontiguousarray(domain, np.float64)
    out = np.zeros((max(cap, 1), 3))
    n = C.c_int(0)
    check(lib.gev_lightcone_vertices(C.byref(g), float(outer), float(inner), dom.ctypes.data_as(C.POINTER(C.c_double)),
                                     out.ctypes.data_as(C.POINTER(C.c_double)), int(cap), C.byref(n)),
          "gev_lightcone_vertices")
    return n.value, out[:min(n.value, cap)].copy()


class IDLog:
    """IDs written near the inner edge of the last cycle's shells (backlog) and
    of this cycle's (prelog); commit() hands the prelog over as the backlog"""

    def __init__(self, ctx):
        self.ctx = ctx
        h = C.c_void_p()
        check(lib.gev_idlog_create(ctx._h, C.byref(h)), "gev_idlog_create")
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib.gev_idlog_destroy(h)
            self._h = None

    def _get(self, which):
        nb, npre = C.c_int64(0), C.c_int64(0)
        check(lib.gev_idlog_sizes(self._h, C.byref(nb), C.byref(npre)), "gev_idlog_sizes")
        n = nb.value if which == 0 else npre.value
        out = np.zeros(max(n, 1), np.int64)
        check(lib.gev_idlog_get(self._h, which, out.ctypes.data_as(C.POINTER(C.c_int64))), "gev_idlog_get")
        return out[:n]

    @property
    def backlog(self):
        return self._get(0)

    @property
    def prelog(self):
        return self._get(1)

    def set_backlog(self, ids):
        a = np.ascontiguousarray(ids, np.int64)
        check(lib.gev_idlog_set_backlog(self._h, a.ctypes.data_as(C.POINTER(C.c_int64)), len(a)),
              "gev_idlog_set_backlog")

    def commit(self):
        check(lib.gev_idlog_commit(self._h), "gev_idlog_commit")


@dataclass
class Output:
    """what write_particles wrote for one lightcone"""
    index: int
    path: str
    written: int
    vertices: int


def write_particles(pcls, lightcones, idlogs, phi, a, tau, dtau, dtau_old, cycle, basename, boxsize, fourpiG,
                    Omega_m, Omega_Lambda, Omega_rad=0.0, Omega_pcl=None, h=0.67, tracer_factor=1, numpcl=None,
                    idlog_mapping=None, species="_cdm"):
    """The Gadget-2 branch of writeLightcones for one species
    (output.hpp:1213, 2094-2160, then the ID-log hand-over 2166-2880): for each
    lightcone whose distance range covers this cycle's look-back shell, find
    the vertices, fill the header and write "<basename>[i]_<cycle>_cdm";
    finally commit every ID log.  boxsize in Mpc/h; a, tau, dtau, dtau_old in
    the code's units.  Returns the outputs written."""
    from . import hconf, particle_horizon
    N = pcls.ctx.N
    dom = box_domain(N)
    numpcl = pcls.n if numpcl is None else numpcl
    Omega_pcl = Omega_m if Omega_pcl is None else Omega_pcl
    mapping = list(range(len(lightcones))) if idlog_mapping is None else list(idlog_mapping)
    outs = []
    for i, lc in enumerate(lightcones):
        d = particle_horizon(1.0 / (1.0 + lc.z), fourpiG, Omega_m, Omega_Lambda, Omega_rad)
        shell = d - tau + 0.5 * dtau_old
        if not (lc.distance[0] > shell and lc.distance[1] <= shell and shell > 0.0):
            continue
        n, verts = vertices(lc, d - tau + (0.5 + IDCHECK_ZONE) * dtau_old, d - tau - 0.5 * dtau, dom)
        if n > MAX_VERTICES:
            raise GevError(f"lightcone {i}: {n} periodic copies meet the shell, more than {MAX_VERTICES}")
        hdr = gev_gadget2_header()
        hdr.num_files = 1
        hdr.Omega0, hdr.OmegaLambda, hdr.HubbleParam = Omega_m, Omega_Lambda, h
        hdr.BoxSize = boxsize / GADGET_LENGTH_CONVERSION
        hdr.time = a
        hdr.redshift = 1.0 / a - 1.0
        hdr.mass[1] = tracer_factor * C_RHO_CRIT * Omega_pcl * boxsize ** 3 / numpcl / GADGET_MASS_CONVERSION
        name = (f"{i}_{cycle:04d}" if len(lightcones) > 1 else f"_{cycle:04d}") + species
        path = basename + name
        dadtau = a * hconf(a, fourpiG, Omega_m, Omega_Lambda, Omega_rad)
        w = pcls.save_gadget2_lightcone(path, hdr, lc, d - tau, dtau, dtau_old, dadtau, verts,
                                        idlogs[mapping[i]], phi, tracer_factor)
        outs.append(Output(i, path if w > 0 else "", w, n))
    for log in {id(x): x for x in idlogs}.values():
        log.commit()
    return outs
