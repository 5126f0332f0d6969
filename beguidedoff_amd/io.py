"""Guide TSV input and the long-format output table.

Input formats and row ids follow beguidedoff/io.py:15-152; the output
columns, their order and the score rounding follow io.py:155-180.
"""
from typing import List, Optional

import pandas as pd

from .base_editing import get_edit_window_from_profile
from .types import GuideInput, OffTarget

OUTPUT_COLUMNS = [
    "guide_rna_sequence", "base_editor", "ot_rank", "ot_chromosome", "ot_position",
    "ot_strand", "ot_sequence", "ot_pam", "n_mismatches", "mismatch_positions", "score",
    "n_offtargets_below_threshold",
]

_EDITING_TO_TYPE = {"ABE": "ABE", "CBE": "CBE"}


def _opt_str(row, key) -> Optional[str]:
    v = row.get(key)
    return str(row.get(key, "")) if pd.notna(v) else None


def _opt_int(row, key) -> Optional[int]:
    v = row.get(key)
    return int(v) if pd.notna(v) else None


def load_guides_from_tsv(input_file: str, presets_path: Optional[str] = None) -> List[GuideInput]:
    """Read guides; ids are ``row_<i>`` (input order), pam_position is 3prime.

    Explicit format: guide_seq, pam, edit_window_start, edit_window_end,
    editor_type [, chromosome, position, on_target_strand, base_editor].
    Profile format: guide_rna_sequence, base_editor [, chromosome, position,
    strand_used]; PAM and window come from the profiles table.
    """
    df = pd.read_csv(input_file, sep="\t")
    guides: List[GuideInput] = []
    if "guide_seq" in df.columns:
        for idx, row in df.iterrows():
            guides.append(GuideInput(
                guide_id=f"row_{idx}",
                guide_seq=str(row["guide_seq"]).upper(),
                pam=str(row["pam"]).upper(),
                pam_position="3prime",
                edit_window_start=int(row["edit_window_start"]),
                edit_window_end=int(row["edit_window_end"]),
                editor_type=str(row["editor_type"]).upper(),
                on_target_chrom=_opt_str(row, "chromosome") or None,
                on_target_pos=_opt_int(row, "position"),
                on_target_strand=_opt_str(row, "on_target_strand"),
                base_editor=str(row.get("base_editor", "")) or None,
            ))
        return guides
    if "guide_rna_sequence" in df.columns and "base_editor" in df.columns:
        if presets_path is None:
            raise ValueError("presets_path is required when using profile-based input "
                             "(guide_rna_sequence + base_editor columns)")
        for idx, row in df.iterrows():
            editor = str(row["base_editor"])
            start, end, pam = get_edit_window_from_profile(editor, presets_path)
            guides.append(GuideInput(
                guide_id=f"row_{idx}",
                guide_seq=str(row["guide_rna_sequence"]).upper(),
                pam=pam.upper(),
                pam_position="3prime",
                edit_window_start=start,
                edit_window_end=end,
                editor_type=_resolve_editor_type(presets_path, editor),
                on_target_chrom=_opt_str(row, "chromosome") or None,
                on_target_pos=_opt_int(row, "position"),
                on_target_strand=_opt_str(row, "strand_used"),
                base_editor=editor,
            ))
        return guides
    raise ValueError("Input file must contain either (guide_seq, pam, ...) or "
                     f"(guide_rna_sequence, base_editor) columns. Found: {list(df.columns)}")


def _resolve_editor_type(profiles_path: str, base_editor: str) -> str:
    prof = pd.read_csv(profiles_path, sep="\t")
    row = prof[prof["name"] == base_editor]
    if row.empty:
        raise ValueError(f"Unknown base editor '{base_editor}' in {profiles_path}")
    editing = str(row.iloc[0]["editing"])
    if editing not in _EDITING_TO_TYPE:
        raise ValueError(f"Base editor '{base_editor}' has unrecognised editing type "
                         f"'{editing}'. Expected one of: {sorted(_EDITING_TO_TYPE)}")
    return _EDITING_TO_TYPE[editing]


def offtargets_to_dataframe(offtargets: List[OffTarget],
                            n_offtargets_below_threshold: int = 0) -> pd.DataFrame:
    rows = [{
        "guide_rna_sequence": ot.guide_seq,
        "base_editor": ot.base_editor,
        "ot_rank": rank,
        "ot_chromosome": ot.chromosome,
        "ot_position": ot.position,
        "ot_strand": ot.strand,
        "ot_sequence": ot.offtarget_seq,
        "ot_pam": ot.genomic_pam,
        "n_mismatches": ot.n_mismatches,
        "mismatch_positions": ",".join(map(str, ot.mismatch_positions)),
        "score": round(ot.score, 6),
        "n_offtargets_below_threshold": n_offtargets_below_threshold,
    } for rank, ot in enumerate(offtargets, start=1)]
    return pd.DataFrame(rows)
