"""Edge-case pairs as #PP 2.0 text: single nucleotides, a pair without any
base pairs, base pairs on one side only, symbols outside ACGU (N), an
identical pair, and a pair whose only base pair has the smallest span the
reference registers (j - i = 3)."""
from locarna_amd.synth import pp_text

HAIRPIN = [(1, 9, 0.9), (2, 8, 0.8)]

CASES = {
    "len1": (pp_text("a", "A", []), pp_text("b", "G", [])),
    "len1_same": (pp_text("a", "C", []), pp_text("b", "C", [])),
    "len1_len5": (pp_text("a", "C", []), pp_text("b", "GGAAC", [(1, 5, 0.6)])),
    "len2_len9": (pp_text("a", "GC", []), pp_text("b", "GGGAAACCC", HAIRPIN)),
    "noarcs": (pp_text("a", "ACGUA", []), pp_text("b", "ACGU", [])),
    "onesided": (pp_text("a", "GGGAAACCC", HAIRPIN), pp_text("b", "ACGUACG", [])),
    "N_symbols": (pp_text("a", "GNNAAUCC", [(1, 8, 0.7)]), pp_text("b", "GGNAACC", [(1, 7, 0.5)])),
    "all_N": (pp_text("a", "NNNNNN", [(1, 6, 0.4)]), pp_text("b", "NNNNN", [(1, 5, 0.3)])),
    "identical": (pp_text("a", "GGGAAACCC", HAIRPIN), pp_text("b", "GGGAAACCC", HAIRPIN)),
    "min_span": (pp_text("a", "GAAC", [(1, 4, 0.8)]), pp_text("b", "GAAAC", [(1, 5, 0.7), (2, 5, 0.2)])),
}
