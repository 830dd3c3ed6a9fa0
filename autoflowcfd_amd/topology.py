"""Face topology on the GPU (libafr.so: afr_face_topology).

The records, codes and FP64 geometry are those of ``mesh.build_topology``
(FaceExtractor.extract_faces_mixed, src/autoflowcfd/grid/mesh_gen/extraction/face_extractor.py:296-376,
and src/autoflowcfd/grid/connectivity/face_connectivity.py:53-231) computed by a device radix sort
of the sorted node triples; the result comes back as host ``FaceRecords`` so that boundary
classification, periodic pairing and boundary groups (host, per boundary face) apply unchanged.
"""

import ctypes

import numpy as np
import torch

from . import _lib
from .mesh import BC_INTERIOR, FaceRecords, check_cells


def build_topology_gpu(mesh, device=None) -> FaceRecords:
    """Deduplicated triangle records of ``mesh`` (prisms first, then tets) built on the GPU."""
    lib = _lib.load()
    if not torch.cuda.is_available():
        raise RuntimeError("build_topology_gpu needs a visible GPU (HIP device)")
    dv = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    P, T = int(mesh.n_prism), int(mesh.n_tet)
    n_nodes = int(mesh.nodes.shape[0])
    if n_nodes >= 2 ** 31:
        raise ValueError("more than 2^31 nodes")
    cap = 8 * P + 4 * T
    if cap == 0:
        raise ValueError("mesh without cells")
    for conn in (mesh.prism_conn, mesh.tet_conn):
        if conn.size and (conn.min() < 0 or conn.max() >= n_nodes):
            raise ValueError("cell connectivity names a node outside the node array")
    check_cells(mesh)
    nodes = torch.as_tensor(np.ascontiguousarray(mesh.nodes, dtype=np.float64)).to(dv)
    pc = torch.as_tensor(np.ascontiguousarray(mesh.prism_conn.reshape(-1, 6), dtype=np.int32)).to(dv)
    tc = torch.as_tensor(np.ascontiguousarray(mesh.tet_conn.reshape(-1, 4), dtype=np.int32)).to(dv)
    nbytes = ctypes.c_size_t(0)
    _lib.check(lib.afr_face_topology_workspace(P, T, ctypes.byref(nbytes)), "afr_face_topology_workspace")
    work = torch.empty((max(int(nbytes.value), 1),), dtype=torch.uint8, device=dv)
    i32 = lambda *s: torch.empty(s, dtype=torch.int32, device=dv)      # noqa: E731
    f64 = lambda *s: torch.empty(s, dtype=torch.float64, device=dv)    # noqa: E731
    out = {"owner": i32(cap), "neighbor": i32(cap), "owner_code": i32(cap), "neighbor_code": i32(cap),
           "node_ids": i32(cap, 3), "normal": f64(cap, 3), "area": f64(cap), "center": f64(cap, 3),
           "n_faces": i32(1), "status": i32(1)}
    p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    d = _lib.AfrTopoDesc()
    d.abi_version, d.n_nodes, d.n_prism, d.n_tet = _lib.ABI_VERSION, n_nodes, P, T
    d.nodes, d.prism_conn, d.tet_conn = p(nodes), p(pc), p(tc)
    d.work, d.work_bytes = p(work), int(work.numel())
    for k, t in out.items():
        setattr(d, k, p(t))
    with torch.cuda.device(dv):
        stream = ctypes.c_void_p(torch.cuda.current_stream(dv).cuda_stream)
        _lib.check(lib.afr_face_topology(ctypes.byref(d), stream), "afr_face_topology")
        torch.cuda.current_stream(dv).synchronize()
    status = int(out["status"].item())
    if status & 1:
        raise RuntimeError("face shared by more than two cells")
    if status & 2:
        raise RuntimeError("ambiguous or missing cube-face match (face_connectivity.py:53-76)")
    F = int(out["n_faces"].item())
    h = {k: v[:F].cpu().numpy() for k, v in out.items() if k not in ("n_faces", "status")}
    i64 = lambda a: a.astype(np.int64)  # noqa: E731
    return FaceRecords(owner=i64(h["owner"]), neighbor=i64(h["neighbor"]), owner_code=i64(h["owner_code"]),
                       neighbor_code=i64(h["neighbor_code"]), normal=h["normal"], area=h["area"],
                       center=h["center"], node_ids=i64(h["node_ids"]), translation=np.zeros((F, 3)),
                       bc=np.full(F, BC_INTERIOR, dtype=np.int64))
