"""The G2 window-combination programs (cusnarks_amd/csrc/coop29.h) as data:
every operand is produced in an earlier round (or is an input slot), no lane
reads a slot that another lane of its round writes, no two lanes write one
slot, no accumulator slot is read after its program has overwritten it, and
each formula (doubling; addition = parts 1 and 2) rewrites all eight
accumulator slots.  A program that broke one of these would still compile."""
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "..", "cusnarks_amd", "csrc", "coop29.h")
ACC = ["H_X0", "H_X1", "H_Y0", "H_Y1", "H_ZZ0", "H_ZZ1", "H_ZZZ0", "H_ZZZ1"]
ADDEND = ["H_QX0", "H_QX1", "H_QY0", "H_QY1", "H_QZZ0", "H_QZZ1", "H_QZZZ0", "H_QZZZ1"]
TWO_OUT = ("C29_DOTN", "C29_DOTMN")  # a second result slot in a4


def _program(src, name):
    i = src.index("__constant__ C29Op %s[] = {" % name)
    ops = re.findall(r"\{(C29_\w+), (\w+), \{([^}]*)\}\}", src[i:src.index("};", i)])
    rounds = [int(x) for x in re.search(r"__constant__ uint8_t %sR\[\] = \{([^}]*)\}" % name, src).group(1).split(",")]
    return ops, rounds


def _io(kind, d, args):
    args = [x.strip() for x in args.split(",")]
    if kind in TWO_OUT:
        return args[:4], [d, args[4]]
    return args, [d]


def _check(src, names):
    avail = set(ACC + ADDEND + ["H_ZERO"])
    written_acc = set()
    for name in names:
        ops, rounds = _program(src, name)
        assert len(ops) == sum(rounds), (name, len(ops), rounds)
        assert max(rounds) <= 64, (name, "more operations in a round than lanes")
        k = 0
        for n in rounds:
            rnd = ops[k:k + n]
            k += n
            reads, writes = set(), []
            for kind, d, a in rnd:
                ins, outs = _io(kind, d, a)
                for x in ins:
                    assert x in avail, (name, kind, d, x, "read before it is produced")
                    assert x not in written_acc, (name, kind, d, x, "old accumulator value already overwritten")
                reads |= set(ins)
                writes += outs
            assert len(writes) == len(set(writes)), (name, "two lanes write one slot")
            assert not set(writes) & reads, (name, "read and written in one round", set(writes) & reads)
            assert "H_ZERO" not in writes and not set(writes) & set(ADDEND), (name, "an input slot is written")
            avail |= set(writes)
            written_acc |= {w for w in writes if w in ACC}
    return written_acc


def test_g2_programs_are_well_formed():
    with open(SRC) as fh:
        src = fh.read()
    assert _check(src, ["kG2Dbl"]) == set(ACC)
    assert _check(src, ["kG2Add1", "kG2Add2"]) == set(ACC)
    # part 1 of the addition leaves the accumulator alone: after it the kernel
    # may still double the accumulator (when the addend equals it)
    assert not _check(src, ["kG2Add1"])
