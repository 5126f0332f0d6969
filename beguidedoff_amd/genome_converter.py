"""Genome artefact (.npz) reader and writers.

The artefact layout is the reference's (beguidedoff/genome_converter.py:15-108,
182-201): ``genome_forward`` and ``genome_rc`` (per-contig reverse complement)
as 2-bit packed uint32, ``chrom_names`` / ``chrom_offsets`` /
``chrom_lengths`` / ``total_bp``, and ``n_mask`` = np.packbits (big bit order)
of uppercase-'N' positions.  FASTA parsing is done here without pysam.
"""
from pathlib import Path
from typing import Dict, Iterator, List, Optional, Tuple

import numpy as np

from .encoding import pack_2bit_to_uint32

_CODE = np.zeros(256, dtype=np.uint8)
for _c, _v in (("A", 0), ("C", 1), ("G", 2), ("T", 3)):
    _CODE[ord(_c)] = _v


def read_fasta(path: str) -> Iterator[Tuple[str, str]]:
    """Yield (name, sequence) records; the name is the header's first word."""
    name, parts = None, []
    with open(path, "r") as fh:
        for line in fh:
            line = line.rstrip("\r\n")
            if line.startswith(">"):
                if name is not None:
                    yield name, "".join(parts)
                name, parts = line[1:].split()[0] if line[1:].split() else "", []
            elif line:
                parts.append(line.strip())
    if name is not None:
        yield name, "".join(parts)


def encode_contigs(contigs: List[Tuple[str, str]], include_rc: bool = True,
                   include_n_mask: bool = True, engine: str = "numpy",
                   device: int = 0) -> Dict[str, np.ndarray]:
    """Contig strings -> artefact arrays (upper-cased; non-ACGT -> A).

    ``engine="hip"`` runs the encoder kernels of libbgo_hip.so on ``device``
    (``bgo_encode_genome``); ``engine="numpy"`` encodes on the host.  Both
    give identical arrays; neither falls back to the other.
    """
    if engine == "hip":
        return _encode_hip(contigs, include_rc, include_n_mask, device)
    if engine != "numpy":
        raise ValueError(f"engine must be 'numpy' or 'hip', got {engine!r}")
    names, offsets, lengths, fwd, rc, nmask = [], [], [], [], [], []
    off = 0
    for name, seq in contigs:
        raw = np.frombuffer(seq.upper().encode("latin-1", "replace"), dtype=np.uint8)
        codes = _CODE[raw]
        names.append(name)
        offsets.append(off)
        lengths.append(raw.size)
        off += raw.size
        fwd.append(codes)
        if include_rc:
            rc.append((3 - codes[::-1]).astype(np.uint8))
        if include_n_mask:
            nmask.append((raw == ord("N")).astype(np.uint8))
    full = np.concatenate(fwd) if fwd else np.zeros(0, np.uint8)
    out = {
        "genome_forward": pack_2bit_to_uint32(full),
        "chrom_names": np.array(names),
        "chrom_offsets": np.array(offsets, dtype=np.int64),
        "chrom_lengths": np.array(lengths, dtype=np.int64),
        "total_bp": np.int64(full.size),
    }
    if include_rc:
        out["genome_rc"] = pack_2bit_to_uint32(np.concatenate(rc) if rc else full)
    if include_n_mask:
        out["n_mask"] = np.packbits(np.concatenate(nmask) if nmask else full)
    return out


def encode_hip(seq: np.ndarray, offsets: np.ndarray, lengths: np.ndarray, device: int = 0,
               include_rc: bool = True, include_n_mask: bool = True):
    """Sequence bytes (uint8, contigs back to back at ``offsets``/``lengths``)
    -> (forward words, reverse-complement words or None, N mask or None,
    kernel ms), by ``bgo_encode_genome`` on ``device``."""
    import ctypes as C

    from . import _lib

    _lib.init(device)
    seq = np.ascontiguousarray(seq, dtype=np.uint8)
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    lengths = np.ascontiguousarray(lengths, dtype=np.int64)
    n = int(seq.size)
    fwd = np.zeros((n + 15) // 16, dtype=np.uint32)
    rc = np.zeros_like(fwd) if include_rc else None
    nmask = np.zeros((n + 7) // 8, dtype=np.uint8) if include_n_mask else None
    ms = C.c_float(0.0)
    _lib.check(_lib.load().bgo_encode_genome(
        int(device), _lib.ptr(seq), n, _lib.ptr(offsets), _lib.ptr(lengths), int(offsets.size),
        _lib.ptr(fwd), _lib.ptr(rc), _lib.ptr(nmask), C.byref(ms)))
    return fwd, rc, nmask, float(ms.value)


def _encode_hip(contigs: List[Tuple[str, str]], include_rc: bool, include_n_mask: bool,
                device: int) -> Dict[str, np.ndarray]:
    raw = [seq.encode("latin-1", "replace") for _, seq in contigs]
    lengths = np.array([len(r) for r in raw], dtype=np.int64)
    offsets = np.zeros(len(raw), dtype=np.int64)
    if len(raw) > 1:
        offsets[1:] = np.cumsum(lengths)[:-1]
    seq = np.frombuffer(b"".join(raw), dtype=np.uint8)
    fwd, rc, nmask, _ = encode_hip(seq, offsets, lengths, device, include_rc, include_n_mask)
    out = {
        "genome_forward": fwd,
        "chrom_names": np.array([name for name, _ in contigs]),
        "chrom_offsets": offsets,
        "chrom_lengths": lengths,
        "total_bp": np.int64(seq.size),
    }
    if include_rc:
        out["genome_rc"] = rc
    if include_n_mask:
        out["n_mask"] = nmask
    return out


def convert(fasta_path: str, output_path: str, include_rc: bool = True,
            include_n_mask: bool = True, chunk_size: int = 100_000_000,
            verbose: bool = True, engine: str = "numpy", device: int = 0) -> None:
    """FASTA -> .npz artefact (beguidedoff/genome_converter.py:15-108); see
    :func:`encode_contigs` for ``engine``."""
    arrays = encode_contigs(list(read_fasta(fasta_path)), include_rc, include_n_mask,
                            engine=engine, device=device)
    np.savez_compressed(output_path, **arrays)
    if verbose:
        # numpy appends ".npz" to a name without it
        written = output_path if str(output_path).endswith(".npz") else f"{output_path}.npz"
        mb = Path(written).stat().st_size / (1024 * 1024)
        print(f"  Total genome: {int(arrays['total_bp']):,} bp; saved {output_path} ({mb:.1f} MB)")


def write_npz(path: str, genome: Dict[str, np.ndarray], compressed: bool = False) -> None:
    keys = ["genome_forward", "genome_rc", "chrom_names", "chrom_offsets", "chrom_lengths",
            "total_bp", "n_mask"]
    arrays = {k: genome[k] for k in keys if k in genome}
    (np.savez_compressed if compressed else np.savez)(path, **arrays)


def load_genome(encoded_path: str) -> dict:
    """Read an artefact (beguidedoff/genome_converter.py:182-201).

    Loaded without pickle: ``chrom_names`` is a unicode array.
    """
    try:
        data = np.load(encoded_path, allow_pickle=False)
    except ValueError as exc:
        raise ValueError(f"{encoded_path} is not an encoded genome artefact (an .npz written by "
                         "encode-genome or build-index)") from exc
    out = {
        "genome_forward": data["genome_forward"],
        "chrom_names": data["chrom_names"],
        "chrom_offsets": data["chrom_offsets"],
        "chrom_lengths": data["chrom_lengths"],
        "total_bp": int(data["total_bp"]),
    }
    if "genome_rc" in data:
        out["genome_rc"] = data["genome_rc"]
    if "n_mask" in data:
        out["n_mask"] = data["n_mask"]
    return out


def verify(encoded_path: str, fasta_path: str, sample_n: int = 10000,
           verbose: bool = True) -> bool:
    """Spot-check 20-nt windows of an artefact against its FASTA."""
    data = load_genome(encoded_path)
    total = int(data["total_bp"])
    if total <= 20:
        return True
    seqs = dict(read_fasta(fasta_path))
    offsets, lengths, names = data["chrom_offsets"], data["chrom_lengths"], data["chrom_names"]
    words = data["genome_forward"]
    rng = np.random.default_rng(42)
    errors = 0
    for pos in rng.integers(0, total - 20, size=min(sample_n, total - 20)):
        k = int(np.searchsorted(offsets, pos, side="right") - 1)
        local = int(pos - offsets[k])
        if local + 20 > int(lengths[k]):
            continue
        ref = seqs[str(names[k])][local:local + 20].upper()
        idx = np.arange(pos, pos + 20)
        got = "".join("ACGT"[(int(words[i >> 4]) >> (2 * (i & 15))) & 3] for i in idx)
        exp = ref.replace("N", "A")
        errors += got != exp
    if verbose:
        print(f"  Verified {sample_n} positions: {errors} errors")
    return errors == 0
