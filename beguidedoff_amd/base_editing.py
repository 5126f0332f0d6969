"""Base-editing window (beguidedoff/base_editing.py:9-91)."""
from typing import List, Tuple

EDITABLE_BASE = {"CBE": "C", "ABE": "A", "CGE": "C"}


def editable_base_in_window(offtarget_seq: str, editor_type: str, edit_window_start: int,
                            edit_window_end: int) -> List[int]:
    """1-based positions in [start, end] holding the editor's target base.

    Indices follow Python slicing of the 0-based range(start - 1, end) over
    the sequence, so a start <= 0 reads from the end of the sequence exactly
    as the reference does.
    """
    target = EDITABLE_BASE.get(editor_type)
    if target is None:
        raise ValueError(f"Unknown editor_type: {editor_type}")
    seq = offtarget_seq.upper()
    return [i + 1 for i in range(edit_window_start - 1, edit_window_end)
            if i < len(seq) and seq[i] == target]


def edit_window_mask(edit_window_start: int, edit_window_end: int, length: int = 20) -> int:
    """Bit i set when sequence index i is inspected by editable_base_in_window."""
    mask = 0
    for i in range(edit_window_start - 1, edit_window_end):
        if i < length and i >= -length:
            mask |= 1 << (i % length)
    return mask


def get_trinucleotide_contexts(seq: str, positions: List[int]) -> List[str]:
    seq = seq.upper()
    out = []
    for pos in positions:
        i = pos - 1
        left = seq[i - 1] if i > 0 else "N"
        right = seq[i + 1] if i < len(seq) - 1 else "N"
        out.append(left + seq[i] + right)
    return out


def has_editable_base(offtarget_seq: str, editor_type: str, edit_window_start: int,
                      edit_window_end: int) -> bool:
    return bool(editable_base_in_window(offtarget_seq, editor_type, edit_window_start,
                                        edit_window_end))


def get_edit_window_from_profile(base_editor: str, profiles_path: str) -> Tuple[int, int, str]:
    """(general_edit_start, general_edit_end, pam) of a profile row."""
    import pandas as pd

    df = pd.read_csv(profiles_path, sep="\t")
    missing = {"name", "general_edit_start", "general_edit_end", "pam"} - set(df.columns)
    if missing:
        raise ValueError(f"Base editor profiles file {profiles_path} is missing required "
                         f"columns: {sorted(missing)}")
    row = df[df["name"] == base_editor]
    if row.empty:
        names = ", ".join(sorted(df["name"].head(5).tolist()))
        more = "..." if len(df) > 5 else ""
        raise ValueError(f"Unknown base editor '{base_editor}' in {profiles_path}. "
                         f"Available profiles include: {names}{more}")
    r = row.iloc[0]
    return int(r["general_edit_start"]), int(r["general_edit_end"]), str(r["pam"])
