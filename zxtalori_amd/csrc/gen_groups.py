"""Writes groups_table.h: the f32 operator quaternions of the point groups.

The operator order and the f32 bits must match the reference's tables
(src/groups.zig:24-201, scalar first, right action q*g) because image codes
(source + n*(2*op+sign)), the fold's tie-breaking and the spacing indices are defined
by them.  The f64 values the reference rounds to f32 are the ones its own tests hold
in tests/fixtures/pair-operators.npz (kept here as tests/golden/pair-operators.npz);
this script rounds them to f32 and prints hex-float literals.  C1 (identity only) is
an extension with no reference table.
"""
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
GROUPS = ("C2", "C3", "C4", "C6", "D2", "D3", "D4", "D6", "T", "O", "I")


def main():
    data = np.load(ROOT / "tests" / "golden" / "pair-operators.npz", allow_pickle=False)
    lines = [
        "// Generated by gen_groups.py from the f64 operator tables of the reference",
        "// (src/groups.zig:24-201); do not edit.  Rows are scalar-first quaternions.",
        "#pragma once",
        "#include <cstdint>",
        "namespace zxt {",
        "constexpr uint32_t kGroupCount = 12;  // ids 0..10 as the reference, 11 = C1",
        "constexpr uint32_t kMaxOrder = 60;",
    ]
    offsets, orders, rows = [], [], []
    for name in GROUPS + ("C1",):
        ops = np.array([[1.0, 0.0, 0.0, 0.0]]) if name == "C1" else data[name]
        offsets.append(len(rows))
        orders.append(len(ops))
        rows.extend(np.asarray(ops, dtype=np.float32).tolist())
    lines.append("constexpr const char* kGroupNames[kGroupCount] = {"
                 + ", ".join(f'"{g}"' for g in GROUPS + ("C1",)) + "};")
    lines.append("constexpr uint32_t kGroupOrder[kGroupCount] = {" + ", ".join(map(str, orders)) + "};")
    lines.append("constexpr uint32_t kGroupOffset[kGroupCount] = {" + ", ".join(map(str, offsets)) + "};")
    lines.append(f"alignas(16) constexpr float kOperators[{len(rows)}][4] = {{")
    for r in rows:
        lines.append("    {" + ", ".join(float(np.float32(v)).hex() + "f" for v in r) + "},")
    lines.append("};")
    lines.append("}  // namespace zxt")
    out = Path(__file__).with_name("groups_table.h")
    out.write_text("\n".join(lines) + "\n")
    print(f"wrote {out} ({len(rows)} operators)")


if __name__ == "__main__":
    main()
