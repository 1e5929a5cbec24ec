"""PDB -> density on the HIP splat kernel, behind the reference's API.

Mirrors pdb2mrc_gpu_ver_fp32_v3.py (density_from_coords p2m:41-98,
yflip_vectorized p2m:101-112, generate_mrc p2m:115-135) and the zip release's
generate_volume (pose-transformed chains).  Volumes are float32 [z, y, x]
tensors on the HIP device.
"""
from __future__ import annotations

import numpy as np

# p2m:13-16: (atomic number, mass) per element symbol; others weigh 0.
element_data = {
    "H": (1.0, 1.00794), "C": (6.0, 12.0107), "N": (7.0, 14.00674),
    "O": (8.0, 15.9994), "P": (15.0, 30.973761), "S": (16.0, 32.066),
}
ELEMENT_Z = {k: v[0] for k, v in element_data.items()}


def atoms_to_atomic_numbers_and_masses(atoms):
    """p2m:19-22."""
    values = np.asarray([element_data.get(str(a).upper(), (0.0, 0.0)) for a in atoms],
                        dtype=np.float64).reshape(-1, 2)
    return values[:, 0], values[:, 1]


def splat_parameters(apix: float, res: float):
    """(width, rp, kn) of p2m:53-58: width = round(3 res / apix), rp = (pi apix / res)^2, kn = (rp/pi)^1.5."""
    width = int(round(res * 3.0 / apix))
    rp = float(np.power(np.pi / (res / apix), 2))
    kn = float(np.power(rp / np.pi, 1.5))
    return width, rp, kn


def _hip_device(device):
    """The HIP device to splat on.  The signatures keep p2m's default
    device='cpu' (p2m:41), but this engine has no CPU path: a call that does
    not name a HIP device (e.g. device='cuda:0') is refused."""
    import torch
    dev = torch.device(device)
    if dev.type != "cuda":
        raise ValueError(f"gisapr_amd splats on HIP devices only: pass device='cuda' or 'cuda:N' (got {dev})")
    if dev.index is None:
        dev = torch.device("cuda", torch.cuda.current_device())
    return dev


def density_from_coords(coordinates, atomic_numbers, boxsize, apix, res, device="cpu",
                        max_chunk_voxels=524288, _initial_volume=None, backend="hip",
                        backend_metrics=None):
    """Radial-exponential density of atoms as a float32 [box, box, box] tensor ([z, y, x]).

    Same validation and rule as p2m:41-98; ``device`` must name a HIP device
    (the default 'cpu' of p2m:41 is refused).  ``_initial_volume`` (the zip
    release's fixed-body reuse) is cloned and the atoms are added on top.
    ``max_chunk_voxels`` is accepted for signature compatibility; the kernel
    needs no chunking.
    """
    import torch
    from . import kernels
    boxsize = int(boxsize)
    apix, res = float(apix), float(res)
    if boxsize < 1 or not np.isfinite(apix) or not np.isfinite(res) or apix <= 0 or res <= 0:
        raise ValueError("boxsize, apix and resolution must be positive and finite")
    width, rp, kn = splat_parameters(apix, res)
    if width < 3:
        raise ValueError("You do not have sufficient sampling for this resolution. Decrease apix.")
    dev = _hip_device(device)
    xyz = torch.as_tensor(coordinates, dtype=torch.float32).reshape(-1, 3).to(dev).contiguous()
    numbers = torch.as_tensor(atomic_numbers, dtype=torch.float32).reshape(-1).to(dev).contiguous()
    if len(xyz) != len(numbers):
        raise ValueError("Coordinate and atom count differ")
    if _initial_volume is not None:
        init = torch.as_tensor(_initial_volume)
        if tuple(init.shape) != (boxsize,) * 3 or init.dtype != torch.float32:
            raise ValueError("_initial_volume must be float32 [box, box, box]")
        volume = init.to(dev).clone()
    else:
        volume = torch.zeros((boxsize,) * 3, dtype=torch.float32, device=dev)
    if len(xyz) == 0:
        return volume
    if not bool(torch.isfinite(xyz).all()):
        raise ValueError("PDB coordinates must be finite")
    kernels.splat(xyz, numbers, boxsize, width, apix, rp, kn, volume)
    if backend_metrics is not None:
        backend_metrics["density_backend"] = "hip"
    return volume


def yflip_vectorized(arr):
    """EMAN y-flip; even boxes keep row zero (p2m:101-112)."""
    ny = arr.shape[1]
    idx = np.arange(ny)
    j1 = np.arange(ny // 2, ny)
    j2 = ny - j1 + (-1 if ny % 2 else 0)
    ok = j1 > j2
    idx[j1[ok]], idx[j2[ok]] = j2[ok], j1[ok]
    if isinstance(arr, np.ndarray):
        return arr[:, idx, :].copy()
    import torch
    return arr.index_select(1, torch.as_tensor(idx, device=arr.device))


def generate_volume(pdb_path, boxsize, apix, res, device="cpu", do_center=False, do_yflip=False,
                    include_hetatm=False, source_pdb=None, chain_id=None, pose=None):
    """Density tensor of a PDB file, or of ``source_pdb`` with ``chain_id`` moved by ``pose``."""
    from .pdb_text import read_pdb, transformed_density_atoms
    if source_pdb is None:
        pdb = read_pdb(pdb_path, include_hetatm=include_hetatm, first_model_only=False)
        coords, elements = pdb.coords, pdb.elements
    else:
        coords, elements = transformed_density_atoms(source_pdb, chain_id, pose, include_hetatm)
    numbers, masses = atoms_to_atomic_numbers_and_masses(elements)
    if do_center and len(coords):
        if masses.sum() <= 0:
            raise ValueError("Cannot center PDB: no supported atoms with nonzero mass")
        coords = coords - np.sum(coords * masses[:, None], axis=0) / masses.sum()
    vol = density_from_coords(coords, numbers, boxsize, apix, res, device=device)
    if do_yflip:
        vol = yflip_vectorized(vol)
    return vol


def generate_mrc(pdb_path, output, boxsize, apix, res, device="cpu", do_center=False, do_yflip=False,
                 include_hetatm=False):
    """Write the density as MRC (returns 0) or return it as a numpy array when output is None."""
    vol = generate_volume(pdb_path, boxsize, apix, res, device, do_center, do_yflip, include_hetatm)
    data = vol.detach().cpu().numpy()
    if output is None:
        return data
    from .mrc import write_mrc
    write_mrc(output, data, voxel_size=float(apix))
    return 0


def create_parser():
    """The pdb2mrc command line (p2m:138-151)."""
    import argparse
    p = argparse.ArgumentParser(description="Read text PDB and generate a 3-D MRC on the HIP device.")
    p.add_argument("--i", required=True, help="Input PDB file")
    p.add_argument("--o", required=True, help="Output MRC file")
    p.add_argument("--box", type=int, default=256)
    p.add_argument("--apix", type=float, default=1.0)
    p.add_argument("--res", type=float, default=2.0)
    p.add_argument("--center", action="store_true")
    p.add_argument("--yflip", action="store_true")
    p.add_argument("--includeHETATM", action="store_true")
    p.add_argument("--gpuid", type=int, default=0)
    p.add_argument("--device", help="HIP device, e.g. cuda:0; overrides --gpuid")
    p.add_argument("--dontWriteMRC", action="store_true")
    return p


def run_pdb2mrc(args):
    """generate_mrc with the parsed flags (p2m:154-162); the density array
    with --dontWriteMRC, else 0 after writing --o."""
    device = getattr(args, "device", None) or f"cuda:{int(args.gpuid)}"
    return generate_mrc(args.i, None if args.dontWriteMRC else args.o, args.box, args.apix, args.res, device=device,
                        do_center=args.center, do_yflip=args.yflip, include_hetatm=args.includeHETATM)


def main(argv=None):
    return run_pdb2mrc(create_parser().parse_args(argv))


if __name__ == "__main__":
    main()
