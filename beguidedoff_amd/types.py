"""API records of the extensive search.

Field names, order and defaults are those of the reference's
``beguidedoff/types.py:6-40`` so that callers can construct them the same way.
"""
from dataclasses import dataclass, field
from typing import List, Optional


@dataclass(frozen=True)
class GuideInput:
    """One guide RNA to search (beguidedoff/types.py:6-20)."""

    guide_id: str
    guide_seq: str            # protospacer, 5'->3'; the first 20 nt are compared
    pam: str                  # 1-6 IUPAC characters, 3' of the protospacer
    pam_position: str         # only "3prime" is accepted by the finder
    edit_window_start: int    # 1-based, from the 5' end of the protospacer
    edit_window_end: int      # 1-based, inclusive
    editor_type: str          # "ABE", "CBE" or "CGE"
    on_target_chrom: Optional[str] = None
    on_target_pos: Optional[int] = None     # 1-based; rightmost base on '-'
    on_target_strand: Optional[str] = None  # "+" or "-"
    base_editor: Optional[str] = None       # editor profile name


@dataclass
class OffTarget:
    """One reported site (beguidedoff/types.py:23-40)."""

    chromosome: str
    position: int             # 1-based; '+': 5' end of the site, '-': rightmost base
    strand: str
    offtarget_seq: str        # 20 nt read from the scanned strand array
    genomic_pam: str
    input_pam: str
    guide_seq: str
    base_editor: str
    n_mismatches: int
    mismatch_positions: List[int]
    score: float
    editable_bases: List[int]
    base_edit_context: List[str]


@dataclass
class GuideResult:
    """Per-guide summary record (beguidedoff/types.py:43-49)."""

    guide_id: str
    total_offtargets: int
    top_offtargets: List[OffTarget] = field(default_factory=list)
