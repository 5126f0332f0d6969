"""Command line of the extensive-mode finder, with the reference's
subcommands and options (beguidedoff/cli.py:75-126):

    python -m beguidedoff_amd.cli encode-genome --fasta G.fa --output G.npz
    python -m beguidedoff_amd.cli search --genome G.npz --guides guides.tsv [--output out.tsv]
    python -m beguidedoff_amd.cli verify --encoded G.npz --fasta G.fa

``search --mode fast`` runs the fast (anchored, score-cutoff) search on the GPU
over an encoded genome: ``build-index --output hg38.fmi`` writes that artefact
under the given name (plus the reference's ``.meta`` and ``.nmask``
sidecars), and ``--index hg38.fmi`` names it.
"""
from __future__ import annotations

import argparse
import logging
import sys
from typing import List, Optional

from . import genome_converter

def _encode_genome(args) -> int:
    """cli.py:9-17: FASTA -> .npz artefact."""
    genome_converter.convert(args.fasta, args.output, include_rc=args.include_rc,
                             include_n_mask=args.include_n_mask, verbose=True,
                             engine=args.engine, device=args.device)
    return 0


def _build_index(args) -> int:
    """cli.py:20-24: the fast mode's index files: the encoded genome artefact
    at --output, with the reference's .meta and .nmask sidecars."""
    from .fast import build_index

    build_index(fasta_path=args.fasta, output_path=args.output, verbose=True)
    return 0


def _search(args) -> int:
    """cli.py:27-59: extensive search of a guide TSV; the table goes to
    --output as TSV, or to stdout."""
    if args.mode == "fast":
        if not args.index:
            raise SystemExit("--index is required for --mode fast")
        from .fast import beguidedofffast

        finder = beguidedofffast(index_path=args.index, search_cutoff=args.search_cutoff,
                                 n_top_offtargets=args.n_top, device=args.device)
    else:
        if not args.genome:
            raise SystemExit("--genome is required for --mode extensive")
        from .finder import beguidedoff

        finder = beguidedoff(genome_path=args.genome, max_mismatches=args.max_mismatches,
                             device=args.device, n_top_offtargets=args.n_top,
                             results_per_guide=args.results_per_guide)
    table = finder.search(input_file=args.guides, return_all=args.return_all,
                          count_score_threshold=args.count_score_threshold,
                          output_file=args.output, include_non_editable=args.include_non_editable,
                          presets_path=args.presets_path, verbose=True)
    if not args.output:
        print(table.to_string(index=False))
    return 0


def _verify(args) -> int:
    """cli.py:62-70: spot-check an artefact against its FASTA."""
    ok = genome_converter.verify(args.encoded, args.fasta, sample_n=args.sample_n, verbose=True)
    return 0 if ok else 1


def build_parser() -> argparse.ArgumentParser:
    p = argparse.ArgumentParser(
        prog="beguidedoff_amd",
        description="BEguidedOFF extensive-mode off-target search on AMD Instinct MI355X")
    sub = p.add_subparsers(dest="command", required=True)

    enc = sub.add_parser("encode-genome", help="Encode a FASTA genome (.npz artefact)")
    enc.add_argument("--fasta", required=True, help="Input FASTA file")
    enc.add_argument("--output", required=True, help="Output .npz path")
    enc.add_argument("--include-rc", action="store_true", default=True,
                     help="Include the reverse complement (always on, as in the reference)")
    enc.add_argument("--include-n-mask", action="store_true", default=True,
                     help="Include the N mask (always on, as in the reference)")
    enc.add_argument("--engine", choices=["numpy", "hip"], default="numpy",
                     help="Encode on the host (numpy) or with the GPU encoder kernels (hip)")
    enc.add_argument("--device", type=int, default=0, help="HIP device for --engine hip")
    enc.set_defaults(func=_encode_genome)

    idx = sub.add_parser("build-index", help="Build the fast mode's index files")
    idx.add_argument("--fasta", required=True, help="Input FASTA file")
    idx.add_argument("--output", required=True, help="Output index path (e.g. hg38.fmi)")
    idx.set_defaults(func=_build_index)

    s = sub.add_parser("search", help="Search for off-targets")
    s.add_argument("--mode", choices=["extensive", "fast"], default="extensive")
    s.add_argument("--genome", default=None, help="Encoded genome .npz")
    s.add_argument("--index", default=None, help="Encoded genome .npz (fast mode)")
    s.add_argument("--guides", required=True, help="Guide TSV input")
    s.add_argument("--max-mismatches", type=int, default=4)
    s.add_argument("--search-cutoff", type=float, default=5.0,
                   help="CCTOP score cutoff T for fast mode (default: 5.0)")
    s.add_argument("--device", default=None, help="HIP device, e.g. cuda:0")
    s.add_argument("--n-top", type=int, default=3, help="Top N off-targets per guide")
    s.add_argument("--results-per-guide", type=int, default=10240,
                   help="Initial per-guide site buffer; overflowing guides are re-scanned "
                        "with a larger one")
    s.add_argument("--return-all", action="store_true")
    s.add_argument("--count-score-threshold", type=float, default=5.0)
    s.add_argument("--output", default=None, help="Output TSV file")
    s.add_argument("--include-non-editable", action="store_true")
    s.add_argument("--presets-path", default=None, help="base_editor_profiles.tsv")
    s.set_defaults(func=_search)

    v = sub.add_parser("verify", help="Verify an encoded genome against its FASTA")
    v.add_argument("--encoded", required=True)
    v.add_argument("--fasta", required=True)
    v.add_argument("--sample-n", type=int, default=10000)
    v.set_defaults(func=_verify)
    return p


def main(argv: Optional[List[str]] = None) -> int:
    args = build_parser().parse_args(argv)
    logging.basicConfig(level=logging.INFO, format="%(message)s")
    return int(args.func(args) or 0)


if __name__ == "__main__":
    sys.exit(main())
