"""Builds an unchanged Fortran application from its TOML description on this library's device entries.

The reference's resident flow (compiler/resident/application.py, directives.py, compiler/scopes/
resident_splice.py) reads an application description, applies its edits, expands its `!$fort`
residency directives and replaces the loop nests of the selected procedures by blocks that keep the
arrays on the device and call generated entries. This module does the same for the procedures that
have hand-written gfx950 entries here: the procedure's executable part becomes a block that binds every
array argument to the application's resident_domain, calls the entry's `run` and falls back to the
original statements when the domain is inactive, an array cannot be kept, or the entry declines.

What is read of the description (compiler/resident/application.py:10-75):
  name, sources, program          the sources in compilation order and the main program
  [variants.<v>] defines          preprocessor definitions of a variant
  [[patch]], [[rewrite]]          edits of the sources before the entries are placed, in that order
  [[selection]] procedures        `module::procedure` patterns (`*` matches every name)
  [[edit]]                        edits after the entries are in place, which may add residency
                                  directives: file, optionally procedure (function = true), one of at,
                                  between, call, module_end (or the whole procedure); action before, after
                                  or replace ({original}); text or template
Directives (compiler/resident/directives.py:85-130): enter [device(d)] [budget(b)] [if(c)], leave,
close, publish [if(c)], observe(arrays).

A procedure the selection names but no entry here implements stays as it is; `plan` reports both.
The generated sources are written to a build directory and compiled with the Fortran compiler that
built the reference's host layer (compiler/runtime/scoped_memory.f90, resident_domain.f90).
"""
from __future__ import annotations

import argparse
import fnmatch
import hashlib
import os
import re
import subprocess
import sys
import tempfile
from dataclasses import dataclass, field
from pathlib import Path

HERE = Path(__file__).resolve().parent
FORTRAN = HERE / "fortran"


@dataclass(frozen=True)
class EntrySpec:
    """A hand-written entry: its interface module (fortran/ftc_entries.f90) and the dummy arguments of
    the procedure it replaces, arrays first (real(8), rank 3), then scalars passed by reference; `extra`
    are further scalars the procedure takes from its module's scope (names the call passes as they are)."""
    module: str
    procedure: str
    interface: str
    arrays: int
    scalars: int
    extra: tuple[str, ...] = ()


ENTRIES = {
    # benchmarks/cdu/src/cdu.f90:113-133 (also compiler/tests/fixtures/stencils/cdu.f90)
    ("momentumadvection", "cdu"): EntrySpec("momentumadvection", "cdu", "fort_shared_momentumadvection_cdu", 4, 6),
    # compiler/tests/fixtures/stencils/cvd.f90:104-125
    ("momentumadvection", "cdv"): EntrySpec("momentumadvection", "cdv", "fort_shared_momentumadvection_cdv", 4, 6),
    # compiler/tests/fixtures/stencils/cdw.f90:104-125
    ("momentumadvection", "cdw"): EntrySpec("momentumadvection", "cdw", "fort_shared_momentumadvection_cdw", 4, 6),
    # compiler/tests/apps/box/boundaries.f90:6-41 and stepper.f90:8-37 (dx, dt, nu are module params)
    ("boundaries", "periodic"): EntrySpec("boundaries", "periodic", "fort_shared_boundaries_periodic", 1, 0),
    ("stepper", "diffuse"): EntrySpec("stepper", "diffuse", "fort_shared_stepper_diffuse", 2, 0, ("dx", "dt", "nu")),
    ("stepper", "store"): EntrySpec("stepper", "store", "fort_shared_stepper_store", 2, 0),
}

DOMAIN_MODULE, DOMAIN, CHECK = "ftc_resident", "ftc_domain", "ftc_require"


def load_description(path: Path) -> dict:
    import tomli

    with open(path, "rb") as f:
        return tomli.load(f)


# ------------------------------------------------------------------------------------------- edits
def _unique_span(text: str, piece: str, what: str) -> tuple[int, int]:
    found = [m.span() for m in re.finditer(re.escape(piece), text)]
    if not found:
        # Tolerant of layout: any run of blanks matches any other.
        pattern = r"\s+".join(re.escape(word) for word in piece.split())
        if piece[:1].isspace():
            pattern = r"[ \t]*" + pattern
        if piece[-1:] == "\n":
            pattern += r"[ \t]*\n"
        found = [m.span() for m in re.finditer(pattern, text)]
    if len(found) != 1:
        raise ValueError(f"{what}: expected one {piece!r}, found {len(found)}")
    return found[0]


def _routine_span(text: str, name: str, function: bool) -> tuple[int, int]:
    """The span of `subroutine name` (or `function name`), header to END line inclusive."""
    kind = "function" if function else "subroutine"
    head = re.compile(rf"^[ \t]*(?=\S)(?!end\b)(?:[\w(),=* \t]*\s)?{kind}\s+{re.escape(name)}\b.*$",
                      re.IGNORECASE | re.MULTILINE)
    found = list(head.finditer(text))
    if len(found) != 1:
        raise ValueError(f"expected one {kind} {name}, found {len(found)}")
    end = re.compile(rf"^[ \t]*end[ \t]*(?:{kind}(?:[ \t]+{re.escape(name)})?)?[ \t]*(?:!.*)?$",
                     re.IGNORECASE | re.MULTILINE).search(text, found[0].end())
    if end is None:
        raise ValueError(f"no END of {kind} {name}")
    stop = text.find("\n", end.end())
    return found[0].start(), len(text) if stop < 0 else stop + 1


def _call_span(text: str, name: str, low: int, high: int, what: str) -> tuple[int, int]:
    """The span of the one CALL statement of `name` in text[low:high], continuation lines included."""
    pattern = re.compile(rf"^[ \t]*call[ \t]+{re.escape(name)}\b", re.IGNORECASE | re.MULTILINE)
    found = [m for m in pattern.finditer(text, low, high)]
    if len(found) != 1:
        raise ValueError(f"{what}: expected one call of {name}, found {len(found)}")
    start = pos = found[0].start()
    while True:
        stop = text.find("\n", pos)
        stop = len(text) if stop < 0 else stop + 1
        if not text[pos:stop].split("!", 1)[0].rstrip().endswith("&"):
            return start, stop
        pos = stop


# An edit is resolved in two stages: a locator turns the description's place (at, between, call,
# module_end, or the whole confining procedure) into a half-open character span [start, end) of the
# file; an action then says what that span becomes. Both stages are tables, so a new locator or
# action is one entry.
_EDIT_KEYS = frozenset({"file", "procedure", "function", "action", "text", "template", "at", "between",
                        "call", "module_end"})
_PLACES = ("at", "between", "call", "module_end")

# What the located piece `piece` becomes, given the edit's inserted text `new`.
_ACTIONS = {
    "before": lambda piece, new: new + piece,
    "after": lambda piece, new: piece + new,
    "replace": lambda piece, new: new.replace("{original}", piece),
}


def _place_at(text, window, value, what):
    lo, _ = window
    a, b = _unique_span(text[window[0]:window[1]], value, what)
    return lo + a, lo + b


def _place_between(text, window, value, what):
    opening, closing = value
    lo, hi = window
    first = lo + _unique_span(text[lo:hi], opening, what)[0]
    return first, first + _unique_span(text[first:hi], closing, what)[1]


def _place_call(text, window, value, what):
    return _call_span(text, value, window[0], window[1], what)


def _place_module_end(text, window, value, what):
    # An empty span just before the END MODULE line of the named module, anywhere in the file.
    hit = re.compile(rf"^[ \t]*end[ \t]*module[ \t]+{re.escape(value)}\b", re.IGNORECASE | re.MULTILINE).search(text)
    if hit is None:
        raise ValueError(f"{what}: module {value} has no end module line to insert at")
    return hit.start(), hit.start()


_LOCATORS = {"at": _place_at, "between": _place_between, "call": _place_call, "module_end": _place_module_end}


def _edit_insert(edit: dict, what: str, root: Path | None) -> str:
    if "template" not in edit:
        return edit.get("text", "")
    if "text" in edit or root is None:
        raise ValueError(f"{what}: give text or a template file, not both (a template is read relative to the "
                         f"description)")
    return (root / edit["template"]).read_text()


def apply_edit(text: str, edit: dict, what: str, root: Path | None = None) -> str:
    """One [[edit]] of the description (compiler/resident/application.py:49-58 define its keys).

    `procedure` (a function when `function = true`) confines the search; the place is one of `at` (a piece
    found exactly once, blanks matched loosely), `between` (from the first piece through the second),
    `call` (the CALL statement of that name, continuation lines included) or `module_end` (just before END
    MODULE of that module); with none of them the place is the whole confining procedure. `action` puts
    `text` (or the file `template` under `root`) before or after the place, or replaces it, `{original}`
    standing for the replaced text."""
    stray = sorted(set(edit) - _EDIT_KEYS)
    if stray:
        raise ValueError(f"{what}: unknown edit keys {stray}")
    action = _ACTIONS.get(edit.get("action"))
    if action is None:
        raise ValueError(f"{what}: action {edit.get('action')!r} is none of before, after or replace")
    places = [key for key in _PLACES if key in edit]
    if len(places) > 1:
        raise ValueError(f"{what}: {' and '.join(places)} given; an edit takes one of at, between, call, "
                         f"module_end")
    new = _edit_insert(edit, what, root)
    window = (0, len(text))
    if "procedure" in edit:
        window = _routine_span(text, edit["procedure"], bool(edit.get("function", False)))
    if places:
        start, end = _LOCATORS[places[0]](text, window, edit[places[0]], what)
    elif "procedure" in edit:
        start, end = window
    else:
        raise ValueError(f"{what}: nothing to place the edit at; it needs at, between, call, module_end or "
                         f"procedure")
    return text[:start] + action(text[start:end], new) + text[end:]


# -------------------------------------------------------------------------------------- directives
# A `!$fort` comment line is read as a sequence of items `word` or `word(argument text)`; the first item
# is the directive, the rest are its clauses. Each directive kind is one row of _KINDS: the clauses it
# accepts, whether it takes an argument list, and the Fortran block it stands for. The block texts are the
# protocol the reference's host layer (compiler/runtime/resident_domain.f90) expects.
DIRECTIVE = re.compile(r"^(?P<indent>[ \t]*)!\$fort[ \t]+(?P<body>.*?)[ \t]*$", re.IGNORECASE)
_WORD = re.compile(r"[ \t]*([A-Za-z_]\w*)[ \t]*")
_ARRAY_NAME = re.compile(r"[A-Za-z_]\w*")
# One field of an array of derived type, `objects(:)%field` (observe only).
_FIELD_OF = re.compile(r"([A-Za-z_]\w*)\s*\(\s*:\s*\)\s*%\s*([A-Za-z_]\w*)")


@dataclass(frozen=True)
class Directive:
    kind: str
    arrays: tuple[str, ...]
    clauses: dict
    line: int


def _items(body: str, line: int) -> list[tuple[str, str | None]]:
    """`word` / `word(...)` items of a directive body; parentheses may nest."""
    items, at = [], 0
    while at < len(body):
        word = _WORD.match(body, at)
        if word is None:
            raise ValueError(f"line {line}: cannot read the residency directive {body!r} from column {at + 1}")
        at, argument = word.end(), None
        if body.startswith("(", at):
            depth, close = 0, at
            for close in range(at, len(body)):
                depth += {"(": 1, ")": -1}.get(body[close], 0)
                if depth == 0:
                    break
            if depth:
                raise ValueError(f"line {line}: the residency directive {body!r} has an unclosed parenthesis")
            argument, at = body[at + 1:close].strip(), close + 1
            while at < len(body) and body[at] in " \t":
                at += 1
        items.append((word.group(1).lower(), argument))
    return items


def _enter_lines(d: Directive) -> list[str]:
    c = d.clauses
    budget = f", int({c['budget']}, c_size_t)" if "budget" in c else ""
    start = f"call {DOMAIN}%begin_epoch(int({c.get('device', '0')}, c_int){budget})"
    return [f"if (.not. {DOMAIN}%active) then", f"  if ({c['if']}) {start}" if "if" in c else f"  {start}", "endif"]


def _guard(d: Directive) -> str:
    return f"{DOMAIN}%active" + (f" .and. ({d.clauses['if']})" if "if" in d.clauses else "")


def _observe_lines(d: Directive, origin: str) -> list[str]:
    # Whole arrays are bound and observed together; a field of an array of derived type reaches its
    # structures through the domain's mirror of that field, when there is one.
    plain = [a for a in d.arrays if _ARRAY_NAME.fullmatch(a)]
    fields = [_FIELD_OF.fullmatch(a).groups() for a in d.arrays if not _ARRAY_NAME.fullmatch(a)]
    body = [f"  ftc_observed({n}) = {DOMAIN}%bind({a}, lbound({a}), .true., origin='{origin}::{a.lower()}')"
            for n, a in enumerate(plain, 1)]
    if plain:
        body.append(f"  call {DOMAIN}%observe(ftc_observed)")
    body += [f"  if (size({o}) > 0) call {DOMAIN}%observe_field(ftc_address({o}(lbound({o},1))%{m}))" for o, m in fields]
    head = [f"integer(c_int64_t) :: ftc_observed({len(plain)})"] if plain else []
    return [*head, f"if ({_guard(d)}) then", *body, "endif"]


# kind: (accepted clauses, takes an argument list, uses iso_c_binding, body lines)
_KINDS = {
    "enter": ({"device", "budget", "if"}, False, True, lambda d, o: _enter_lines(d)),
    "leave": (set(), False, False, lambda d, o: [f"if ({DOMAIN}%active) call {DOMAIN}%end_epoch()"]),
    "close": (set(), False, False, lambda d, o: [f"call {DOMAIN}%close()"]),
    "publish": ({"if"}, False, False, lambda d, o: [f"if ({_guard(d)}) call {DOMAIN}%publish()"]),
    "observe": ({"if"}, True, True, _observe_lines),
}


def parse_directive(body: str, line: int) -> Directive:
    """The directive a `!$fort` body states (compiler/resident/directives.py:1-36 list its grammar)."""
    items = _items(body, line)
    if not items:
        raise ValueError(f"line {line}: the residency directive is empty")
    (kind, argument), rest = items[0], items[1:]
    if kind not in _KINDS:
        raise ValueError(f"line {line}: {kind!r} is not a residency directive here (one of {', '.join(_KINDS)})")
    accepted, listed, _, _ = _KINDS[kind]
    clauses = dict(rest)
    bad = [w for w, v in rest if w not in accepted or v is None] + ([] if len(clauses) == len(rest) else ["repeated"])
    if bad:
        raise ValueError(f"line {line}: {kind} accepts {sorted(accepted) or 'no clauses'}; "
                         f"unsupported clause {bad[0]!r} in {body!r}")
    if listed and argument is None:
        raise ValueError(f"line {line}: {kind} needs the list of arrays it brings back")
    if not listed and argument is not None:
        raise ValueError(f"line {line}: {kind} takes no argument list")
    arrays = tuple(a.strip() for a in argument.split(",")) if argument else ()
    malformed = [a for a in arrays if not _ARRAY_NAME.fullmatch(a) and not (kind == "observe" and _FIELD_OF.fullmatch(a))]
    if malformed:
        raise ValueError(f"line {line}: {kind} lists whole arrays by name; {malformed[0]!r} is not an array name")
    return Directive(kind, arrays, clauses, line)


def directive_block(d: Directive, indent: str, origin: str) -> str:
    """The Fortran block that replaces the directive line: a BLOCK that uses the domain object."""
    _, _, c_binding, body = _KINDS[d.kind]
    head = ["block", *(["  use, intrinsic :: iso_c_binding"] if c_binding else []),
            f"  use {DOMAIN_MODULE}, only: {DOMAIN}" + (", ftc_address" if any("%" in a for a in d.arrays) else "")]
    return "".join(f"{indent}{text}\n" for text in (*head, *(f"  {b}" for b in body(d, origin)), "end block"))


def expand_directives(text: str, origin: str = "source") -> tuple[str, list[dict]]:
    """`text` with every `!$fort` line replaced by its block, and one record per directive."""
    pieces, records = [], []
    for number, line in enumerate(text.splitlines(keepends=True), 1):
        found = DIRECTIVE.match(line.rstrip("\r\n"))
        if found is None:
            pieces.append(line)
            continue
        d = parse_directive(found.group("body"), number)
        pieces.append(directive_block(d, found.group("indent"), origin))
        records.append({"line": number, "directive": d.kind, "arrays": list(d.arrays), "clauses": d.clauses})
    return "".join(pieces), records


# ---------------------------------------------------------------------------------- procedure splice
_MODULE = re.compile(r"^\s*module\s+(\w+)\s*(?:!.*)?$", re.IGNORECASE | re.MULTILINE)
_END_MODULE = re.compile(r"^\s*end\s*module\b", re.IGNORECASE | re.MULTILINE)
_SPEC = re.compile(r"^\s*(real|integer|logical|character|complex|double\s*precision|type\s*\(|class\s*\(|use\b|"
                   r"implicit\b|dimension\b|parameter\b|intent\b|external\b|intrinsic\b|save\b|contiguous\b|"
                   r"target\b|optional\b|pointer\b|allocatable\b|import\b|procedure\s*\()", re.IGNORECASE)


def _modules(text: str) -> list[tuple[str, int, int]]:
    spans = []
    for m in _MODULE.finditer(text):
        if m.group(1).lower() == "procedure":
            continue
        end = _END_MODULE.search(text, m.end())
        spans.append((m.group(1).lower(), m.start(), end.start() if end else len(text)))
    return spans


def _logical_lines(lines: list[str], start: int) -> int:
    """Index of the line after the (possibly continued) statement that begins at `start`."""
    i = start
    while i < len(lines) and lines[i].split("!", 1)[0].rstrip().endswith("&"):
        i += 1
    return i + 1


_HEADER = r"^\s*(?!end\b)(?:[\w(),=*]+\s+)*?(subroutine|function)\s+{name}\s*\("


def find_procedure(text: str, module: str, name: str) -> tuple[int, int, int, list[str]] | None:
    """(header line, first executable line, END line, dummy names) of `subroutine name` (or `function
    name`) in `module`, as line indices of text.splitlines(keepends=True); None when it is not there."""
    lines = text.splitlines(keepends=True)
    offsets = [0]
    for line in lines:
        offsets.append(offsets[-1] + len(line))
    for mod, lo, hi in _modules(text):
        if mod != module:
            continue
        header = re.compile(_HEADER.format(name=re.escape(name)), re.IGNORECASE)
        for i, line in enumerate(lines):
            if not (lo <= offsets[i] < hi) or not header.match(line):
                continue
            stop = _logical_lines(lines, i)
            statement = "".join(l.split("!", 1)[0].rstrip().rstrip("&") for l in lines[i:stop])
            opened = header.match(statement).end()
            closed = statement.index(")", opened)
            dummies = [d.strip() for d in statement[opened:closed].split(",") if d.strip()]
            j = stop
            while j < len(lines):
                # An OpenMP directive belongs to the statement after it: the executable part starts there.
                if re.match(r"\s*!\$omp\b", lines[j], re.IGNORECASE):
                    break
                code = lines[j].split("!", 1)[0].strip()
                if code and not _SPEC.match(code):
                    break
                j = _logical_lines(lines, j) if code else j + 1
            kind = header.match(line).group(1).lower()
            end = re.compile(rf"^\s*end\s*(?:{kind}(?:\s+{name})?)?\s*(?:!.*)?$", re.IGNORECASE)
            k = j
            while k < len(lines) and not end.match(lines[k]):
                k += 1
            if k == len(lines):
                raise ValueError(f"no END of {kind} {name}")
            return i, j, k, dummies
    return None


def splice_procedure(text: str, spec: EntrySpec, origin: str) -> str:
    """The procedure's executable part wrapped in the resident block that calls the entry."""
    found = find_procedure(text, spec.module, spec.procedure)
    if found is None:
        raise ValueError(f"{spec.module}::{spec.procedure} not found")
    _, first, end, dummies = found
    if len(dummies) != spec.arrays + spec.scalars:
        raise ValueError(f"{spec.module}::{spec.procedure} has {len(dummies)} arguments; its entry takes "
                         f"{spec.arrays} arrays and {spec.scalars} scalars")
    lines = text.splitlines(keepends=True)
    arrays, scalars = dummies[:spec.arrays], dummies[spec.arrays:] + list(spec.extra)
    ind = "    "
    n = len(arrays)
    arguments = [f"ftc_handles({i})" for i in range(1, n + 1)] + ([", ".join(scalars)] if scalars else [])
    block = [
        "block",
        "  use, intrinsic :: iso_c_binding",
        "  use fort_scoped_memory, only: FORT_SCOPE_GPU, FORT_SCOPE_DECLINED, FORT_SCOPE_RESOURCE",
        f"  use {DOMAIN_MODULE}, only: {DOMAIN}, {CHECK}",
        f"  use {spec.interface}, only: ftc_run => run",
        f"  integer(c_int64_t) :: ftc_handles({n})",
        "  integer(c_int) :: ftc_status",
        "  logical :: ftc_resident_call",
        f"  ftc_resident_call = {DOMAIN}%active",
        "  if (ftc_resident_call) then",
        *(f"    ftc_handles({i}) = {DOMAIN}%bind_real64({a}, lbound({a}), .true., &\n"
          f"{ind}      origin='{origin}::{a.lower()}', tolerant=.true.)" for i, a in enumerate(arrays, 1)),
        # An array that cannot be kept as it is bound here has no handle: the statements stay on the host.
        "    ftc_resident_call = all(ftc_handles /= 0_c_int64_t)",
        "  endif",
        "  if (ftc_resident_call) then",
        f"    ftc_status = ftc_run({DOMAIN}%context, FORT_SCOPE_GPU, &\n"
        + f", &\n".join(f"{ind}      {a}" for a in arguments) + ")",
        # An entry declines what it does not take; one whose arrays do not fit the device budget together
        # is refused (RESOURCE): both run the procedure's own statements on the host
        # (compiler/emission/cuda/scoped.py:472-478, scoped_runtime.cu:1936-1943).
        "    ftc_resident_call = ftc_status /= FORT_SCOPE_DECLINED .and. ftc_status /= FORT_SCOPE_RESOURCE",
        f"    if (ftc_resident_call) call {CHECK}(ftc_status)",
        "  endif",
        "  if (.not. ftc_resident_call) then",
        f"    if ({DOMAIN}%active) call {DOMAIN}%publish()",
    ]
    head = "".join(ind + b + "\n" for b in block)
    tail = "".join(ind + b + "\n" for b in ["  endif", "end block"])
    return "".join(lines[:first]) + head + "".join(lines[first:end]) + tail + "".join(lines[end:])


# ------------------------------------------------------------------------------------------ loop nests
# The reference's compiler cuts the loop nests of a selected main program into generated entries named by a
# hash of the nest (compiler/emission/cuda/scoped.py:61). Here a nest is spliced onto a hand-written entry
# when the hash of its normalised text is one this table lists; the program's other statements that touch
# the nests' arrays run on the host inside host-access blocks.
@dataclass(frozen=True)
class NestEntry:
    """A hand-written entry for one loop nest or array assignment: its interface module (fortran/
    ftc_entries.f90), the nest's arrays in the entry's order and its scalars, passed by reference under the
    names the nest uses."""
    interface: str
    arrays: tuple[str, ...]
    scalars: tuple[str, ...]


NEST_ENTRIES = {
    # compiler/tests/apps/relax/relax.f90:29-35, 36-43 and 44
    "e86ab74e21c687e9": NestEntry("fort_shared_relax_sweep", ("v", "u"), ("n",)),
    "7f1afb921f136343": NestEntry("fort_shared_relax_change", ("v", "u"), ("n", "change")),
    "7c2c9121a204dae3": NestEntry("fort_shared_relax_copy", ("u", "v"), ("n",)),
}


def _code(line: str) -> str:
    """The line without its comment (a `!` outside character constants)."""
    quote = None
    for pos, ch in enumerate(line):
        if quote:
            quote = None if ch == quote else quote
        elif ch in "'\"":
            quote = ch
        elif ch == "!":
            return line[:pos]
    return line


def _statement(lines: list[str], i: int) -> tuple[str, int]:
    """The code of the statement that begins at line i, continuations joined, and the line after it."""
    parts = []
    while True:
        code = _code(lines[i]).strip()
        more = code.endswith("&")
        parts.append(code.rstrip("&").lstrip("&").strip())
        i += 1
        if not more or i == len(lines):
            return " ".join(parts), i


def nest_key(lines: list[str]) -> str:
    """The key of a loop nest or a statement: a hash of its code with comments, directives, blanks and
    letter case taken out, so that layout and OpenMP clauses do not change it."""
    parts, i = [], 0
    while i < len(lines):
        if not _code(lines[i]).strip():
            i += 1
            continue
        text, i = _statement(lines, i)
        parts.append(re.sub(r"\s+", "", text.lower()))
    return hashlib.sha256(";".join(parts).encode()).hexdigest()[:16]


@dataclass
class _Item:
    """A statement or construct of an executable part. start..stop spans it with the OpenMP directives
    that belong to it (the ones before it, and a closing one after it); first is its own first line."""
    start: int
    stop: int
    first: int
    kind: str  # "statement", "construct", "middle" (ELSE, CASE ...) or "directive" (!$fort)
    body: list = field(default_factory=list)


def _balanced_end(text: str, open_at: int) -> int:
    depth = 0
    for pos in range(open_at, len(text)):
        depth += (text[pos] == "(") - (text[pos] == ")")
        if depth == 0:
            return pos
    return len(text)


def _opens(code: str) -> bool:
    s = re.sub(r"^\w+\s*:\s*(?=\w)", "", code)  # a construct name
    if re.match(r"do\b", s) or re.match(r"(select\s*(case|type|rank)|associate|critical)\b", s):
        return True
    if re.fullmatch(r"block", s):
        return True
    for word in ("if", "where", "forall"):
        m = re.match(rf"{word}\s*\(", s)
        if m:
            rest = s[_balanced_end(s, m.end() - 1) + 1:].strip()
            return rest == "then" if word == "if" else rest == ""
    return False


_CLOSES = re.compile(r"end\s*(do|if|select|block|where|forall|associate|critical)\b")
_MIDDLE = re.compile(r"(else\b|elseif\b|elsewhere\b|case\b|class\s+(is|default)\b|type\s+is\b|rank\b)")


def _parse(lines: list[str], lo: int, hi: int) -> list[_Item]:
    """The statements and constructs of lines[lo:hi], constructs holding their bodies."""
    root: list[_Item] = []
    stack: list[list[_Item]] = [root]
    opened: list[_Item] = []
    pending = None  # first line of the OpenMP directives that precede the next statement
    i = lo
    while i < hi:
        s = lines[i].strip().lower()
        if s.startswith("!$omp"):
            j = i + 1
            while lines[j - 1].rstrip().endswith("&") and j < hi:
                j += 1
            if re.match(r"!\$omp\s+end\b", s):
                if stack[-1]:
                    stack[-1][-1].stop = j
            elif pending is None:
                pending = i
            i = j
            continue
        if s.startswith("!$fort"):
            stack[-1].append(_Item(i, i + 1, i, "directive"))
            i += 1
            continue
        if not _code(lines[i]).strip():
            i += 1
            continue
        text, j = _statement(lines, i)
        code = text.lower()
        start, pending = (i if pending is None else pending), None
        if _CLOSES.match(code) or re.fullmatch(r"end\s*(do|if)", code):
            if not opened:
                raise ValueError(f"line {i + 1}: END without its construct")
            opened.pop().stop = j
            stack.pop()
        elif _MIDDLE.match(code) and _assignment_at(code) < 0:
            stack[-1].append(_Item(i, j, i, "middle"))
        else:
            item = _Item(start, j, i, "construct" if _opens(code) else "statement")
            stack[-1].append(item)
            if item.kind == "construct":
                opened.append(item)
                stack.append(item.body)
        i = j
    if opened:
        raise ValueError(f"line {opened[-1].first + 1}: construct without its END")
    return root


def _assignment_at(code: str) -> int:
    """The position of the statement's assignment `=` (outside parentheses, not part of ==, /=, <=, >= or
    =>), or -1."""
    depth = 0
    for pos, ch in enumerate(code):
        depth += (ch == "(") - (ch == ")")
        if (ch == "=" and depth == 0 and code[pos - 1:pos] not in ("=", "/", "<", ">")
                and code[pos + 1:pos + 2] not in ("=", ">")):
            return pos
    return -1


def _written(code: str, names: set[str]) -> set[str]:
    """The arrays of `names` a statement may change: on the left of its assignment, in the argument list of
    a CALL, or in a READ, ALLOCATE or DEALLOCATE."""
    if re.match(r"(call|read|allocate|deallocate)\b", code):
        return {w for w in re.findall(r"[a-z_]\w*", code) if w in names}
    pos = _assignment_at(code)
    return {w for w in re.findall(r"[a-z_]\w*", code[:pos]) if w in names} if pos >= 0 else set()


def _accesses(lines: list[str], item: _Item, names: set[str]) -> tuple[list[str], list[str]]:
    """The arrays of `names` an item reads only, and the ones it may change, in order of appearance."""
    seen: list[str] = []
    written: set[str] = set()
    i = item.first
    while i < item.stop:
        if lines[i].lstrip().lower().startswith("!$") or not _code(lines[i]).strip():
            i += 1
            continue
        text, i = _statement(lines, i)
        code = text.lower()
        for w in re.findall(r"[a-z_]\w*", code):
            if w in names and w not in seen:
                seen.append(w)
        written |= _written(code, names)
    return [a for a in seen if a not in written], [a for a in seen if a in written]


def _entry_block(interface: str, arrays: list[str], scalars: list[str], origin: str) -> list[str]:
    """The head of the resident block that calls an entry; the original statements follow it, then
    `endif` and `end block`."""
    ind = "    "
    n = len(arrays)
    arguments = [f"ftc_handles({i})" for i in range(1, n + 1)] + ([", ".join(scalars)] if scalars else [])
    return [
        "block",
        "  use, intrinsic :: iso_c_binding",
        "  use fort_scoped_memory, only: FORT_SCOPE_GPU, FORT_SCOPE_DECLINED, FORT_SCOPE_RESOURCE",
        f"  use {DOMAIN_MODULE}, only: {DOMAIN}, {CHECK}",
        f"  use {interface}, only: ftc_run => run",
        f"  integer(c_int64_t) :: ftc_handles({n})",
        "  integer(c_int) :: ftc_status",
        "  logical :: ftc_resident_call",
        f"  ftc_resident_call = {DOMAIN}%active",
        "  if (ftc_resident_call) then",
        *(f"    ftc_handles({i}) = {DOMAIN}%bind_real64({a}, lbound({a}), .true., &\n"
          f"{ind}      origin='{origin}::{a.lower()}', tolerant=.true.)" for i, a in enumerate(arrays, 1)),
        # An array that cannot be kept as it is bound here has no handle: the statements stay on the host.
        "    ftc_resident_call = all(ftc_handles /= 0_c_int64_t)",
        "  endif",
        "  if (ftc_resident_call) then",
        f"    ftc_status = ftc_run({DOMAIN}%context, FORT_SCOPE_GPU, &\n"
        + ", &\n".join(f"{ind}      {a}" for a in arguments) + ")",
        # An entry declines what it does not take; one whose arrays do not fit the device budget together
        # is refused (RESOURCE): both run the procedure's own statements on the host
        # (compiler/emission/cuda/scoped.py:472-478, scoped_runtime.cu:1936-1943).
        "    ftc_resident_call = ftc_status /= FORT_SCOPE_DECLINED .and. ftc_status /= FORT_SCOPE_RESOURCE",
        f"    if (ftc_resident_call) call {CHECK}(ftc_status)",
        "  endif",
        "  if (.not. ftc_resident_call) then",
        f"    if ({DOMAIN}%active) call {DOMAIN}%publish()",
    ]


def _host_block(reads: list[str], writes: list[str]) -> tuple[list[str], list[str]]:
    """Head and tail of a host-access block: the arrays read are brought up to date and stay valid on the
    device; the ones changed are the host's until the tail."""
    held = lambda items: "[" + ", ".join(f"{DOMAIN}%held({a})" for a in items) + "]"
    head = ["block", f"  use {DOMAIN_MODULE}, only: {DOMAIN}", "  logical :: ftc_host", f"  ftc_host = {DOMAIN}%active"]
    head += [f"  if (ftc_host) call {DOMAIN}%observe({held(reads)})"] if reads else []
    head += [f"  if (ftc_host) call {DOMAIN}%host_modify({held(writes)})"] if writes else []
    tail = ([f"  if (ftc_host) call {DOMAIN}%host_done({held(writes)})"] if writes else []) + ["end block"]
    return head, tail


_PROGRAM = re.compile(r"^[ \t]*program[ \t]+(\w+)", re.IGNORECASE | re.MULTILINE)


def find_program(text: str, name: str) -> tuple[int, int] | None:
    """(first executable line, END PROGRAM line) of `program name`, as line indices; None when absent."""
    lines = text.splitlines(keepends=True)
    for i, line in enumerate(lines):
        m = _PROGRAM.match(line)
        if not m or m.group(1).lower() != name:
            continue
        j = _logical_lines(lines, i)
        while j < len(lines):
            code = _code(lines[j]).strip()
            if code and not _SPEC.match(code) and not lines[j].lstrip().lower().startswith("!$"):
                break
            j = _logical_lines(lines, j) if code else j + 1
        end = re.compile(rf"^\s*end\s*(?:program(?:\s+{name})?)?\s*(?:!.*)?$", re.IGNORECASE)
        k = j
        while k < len(lines) and not end.match(lines[k]):
            k += 1
        if k == len(lines):
            raise ValueError(f"no END of program {name}")
        return j, k
    return None


def splice_nests(text: str, program: str) -> tuple[str, list[dict]]:
    """`text` with the nests of `program` that have entries spliced onto them, and the other statements
    between its `!$fort enter` and `!$fort leave` that touch their arrays wrapped in host-access blocks;
    and one record per spliced nest."""
    found = find_program(text, program)
    if found is None:
        raise ValueError(f"program {program} not found")
    first, end = found
    lines = text.splitlines(keepends=True)
    items = _parse(lines, first, end)
    spliced: list[tuple[_Item, NestEntry]] = []

    def walk(body):
        for item in body:
            if item.kind in ("statement", "construct"):
                entry = NEST_ENTRIES.get(nest_key(lines[item.first:item.stop]))
                if entry:
                    spliced.append((item, entry))
                    continue
            walk(item.body)

    walk(items)
    device = {a for _, e in spliced for a in e.arrays}
    inside = {id(item) for item, _ in spliced}
    contains = lambda item: id(item) in inside or any(contains(c) for c in item.body)
    marks = [i for i, l in enumerate(lines[first:end], first) if re.match(r"\s*!\$fort\s+(enter|leave)\b", l, re.I)]
    lo, hi = (marks[0], marks[-1]) if len(marks) >= 2 else (first, end)
    edits: list[tuple[int, int, list[str], list[str]]] = []
    for item, entry in spliced:
        edits.append((item.start, item.stop,
                      _entry_block(entry.interface, list(entry.arrays), list(entry.scalars), program),
                      ["  endif", "end block"]))

    def touches(i: int) -> bool:
        return any(w in device for w in re.findall(r"[a-z_]\w*", _statement(lines, i)[0].lower()))

    def guard(body):
        for item in body:
            if item.kind not in ("statement", "construct") or id(item) in inside or not lo < item.first < hi:
                continue
            # A construct with a spliced nest inside, or a branch whose conditions do not name the arrays,
            # is guarded statement by statement, so that the arrays move only where the host uses them; a
            # loop is guarded whole, so that they move once and not per iteration.
            branch = item.kind == "construct" and not re.match(r"(\w+\s*:\s*)?(do|forall|where)\b",
                                                               _statement(lines, item.first)[0].lower())
            heads = [item.first] + [m.first for m in item.body if m.kind == "middle"]
            if contains(item) or (branch and not any(touches(i) for i in heads)):
                guard(item.body)
                continue
            reads, writes = _accesses(lines, item, device)
            if reads or writes:
                edits.append((item.start, item.stop, *_host_block(reads, writes)))

    guard(items)
    for start, stop, head, tail in sorted(edits, key=lambda e: e[0], reverse=True):
        ind = re.match(r"[ \t]*", lines[start]).group(0)
        lines[start:stop] = ([ind + h + "\n" for h in head] + lines[start:stop] + [ind + t + "\n" for t in tail])
    records = [{"program": program, "line": item.first + 1, "entry": entry.interface} for item, entry in spliced]
    return "".join(lines), records


_INTENT = re.compile(r"intent\s*\(\s*(in|out|inout|in\s+out)\s*\)", re.IGNORECASE)


def _dummy_arrays(lines: list[str], header: int, first: int, dummies: list[str]) -> dict[str, str]:
    """The array dummies of a procedure with their intent ('in', 'out', 'inout'; 'inout' when none is
    declared), from its specification statements."""
    names = {d.lower() for d in dummies}
    found: dict[str, str] = {}
    i = _logical_lines(lines, header)
    while i < first:
        stop = _logical_lines(lines, i)
        statement = " ".join(l.split("!", 1)[0].strip().rstrip("&") for l in lines[i:stop])
        i = stop
        if "::" not in statement:
            continue
        attrs, entities = statement.split("::", 1)
        intent = _INTENT.search(attrs)
        mode = intent.group(1).lower().replace(" ", "") if intent else "inout"
        dimensioned = re.search(r"\bdimension\s*\(", attrs, re.IGNORECASE) is not None
        depth, start, parts = 0, 0, []
        for pos, ch in enumerate(entities + ","):
            depth += (ch == "(") - (ch == ")")
            if ch == "," and depth == 0:
                parts.append(entities[start:pos].strip())
                start = pos + 1
        for part in parts:
            m = re.match(r"([A-Za-z_]\w*)\s*(\()?", part)
            if m and m.group(1).lower() in names and (m.group(2) or dimensioned):
                found[m.group(1)] = mode
    return found


def host_access(text: str, module: str, name: str) -> str:
    """A selected procedure without an entry stays on the host: its executable part is wrapped in the
    block that hands it the arrays its caller may keep on the device (compiler/scopes/resident_splice.py:
    228-313 generates the same protocol): the arrays it only reads are brought up to date and stay on
    the device, the ones it may change are handed to the host until it is through. Statements that can
    leave early (RETURN) make the block publish everything instead."""
    found = find_procedure(text, module, name)
    if found is None:
        raise ValueError(f"{module}::{name} not found")
    header, first, end, dummies = found
    lines = text.splitlines(keepends=True)
    arrays = _dummy_arrays(lines, header, first, dummies)
    if not arrays:
        return text
    reads = [a for a, mode in arrays.items() if mode == "in"]
    writes = [a for a, mode in arrays.items() if mode != "in"]
    ind = "    "
    held = lambda items: "[" + ", ".join(f"{DOMAIN}%held({a})" for a in items) + "]"
    if any(re.match(r"^\s*return\b", l, re.IGNORECASE) for l in lines[first:end]):
        head = ["block", f"  use {DOMAIN_MODULE}, only: {DOMAIN}", f"  if ({DOMAIN}%active) call {DOMAIN}%publish()"]
        tail = ["end block"]
    else:
        head = ["block", f"  use {DOMAIN_MODULE}, only: {DOMAIN}", "  logical :: ftc_host", f"  ftc_host = {DOMAIN}%active"]
        head += [f"  if (ftc_host) call {DOMAIN}%observe({held(reads)})"] if reads else []
        head += [f"  if (ftc_host) call {DOMAIN}%host_modify({held(writes)})"] if writes else []
        tail = ([f"  if (ftc_host) call {DOMAIN}%host_done({held(writes)})"] if writes else []) + ["end block"]
    return ("".join(lines[:first]) + "".join(ind + h + "\n" for h in head) + "".join(lines[first:end])
            + "".join(ind + h + "\n" for h in tail) + "".join(lines[end:]))


def selected(description: dict, module: str, procedure: str) -> bool:
    for table in description.get("selection", ()):
        for pattern in table.get("procedures", ()):
            pm, _, pp = pattern.lower().partition("::")
            if fnmatch.fnmatch(module, pm) and fnmatch.fnmatch(procedure, pp or "*"):
                return True
    return False


def plan(description: dict, texts: dict[str, str]) -> dict:
    """Which selected procedures are spliced onto entries and which stay on the host."""
    spliced, native = [], []
    for name, text in texts.items():
        for module, lo, hi in _modules(text):
            for m in re.finditer(r"^[ \t]*(?!end\b)(?:[\w(),=*]+[ \t]+)*?(?:subroutine|function)[ \t]+(\w+)[ \t]*\(",
                                 text[lo:hi], re.IGNORECASE | re.MULTILINE):
                proc = m.group(1).lower()
                if not selected(description, module, proc):
                    continue
                (spliced if (module, proc) in ENTRIES else native).append((name, f"{module}::{proc}"))
    # A selected main program ("$program::name") has its loop nests spliced onto the entries listed for them.
    programs = [(name, m.group(1).lower()) for name, text in texts.items() for m in _PROGRAM.finditer(text)
                if selected(description, "$program", m.group(1).lower())]
    return {"spliced": spliced, "native": native, "programs": programs}


def generate(toml_path: Path, out_dir: Path, source_root: Path | None = None) -> tuple[list[Path], dict]:
    """Writes the application's sources, edited, spliced and with directives expanded, into out_dir;
    returns them in compilation order and the plan. Sources are found relative to `source_root`, by
    default the description's directory."""
    description = load_description(toml_path)
    root = source_root or toml_path.parent
    sources = [root / s for s in description["sources"]]
    texts = {p.name: p.read_text() for p in sources}

    def edits(kind):
        for number, edit in enumerate(description.get(kind, ()), 1):
            if edit.get("file") not in texts:
                raise ValueError(f"{kind} {number} names a file that is not a source: {edit.get('file')}")
            texts[edit["file"]] = apply_edit(texts[edit["file"]], edit, f"{kind} {number} ({edit['file']})",
                                             toml_path.parent)

    # The order of compiler/resident/application.py:45-47: patches of the snapshot, rewrites of the sources
    # before the compiler's passes, then, with the entry calls in place, the edits.
    edits("patch")
    edits("rewrite")
    chosen = plan(description, texts)
    for name, qualified in chosen["spliced"]:
        module, proc = qualified.split("::")
        texts[name] = splice_procedure(texts[name], ENTRIES[module, proc], qualified)
    for name, qualified in chosen["native"]:
        module, proc = qualified.split("::")
        texts[name] = host_access(texts[name], module, proc)
    chosen["nests"] = []
    for name, program in chosen["programs"]:
        texts[name], records = splice_nests(texts[name], program)
        chosen["nests"] += [{"file": name, **r} for r in records]
    edits("edit")
    # The compiler's own loop rewrites reshape nests it generates entries for; hand-written entries have none.
    chosen["ignored_rewrites"] = {f: list(v) for f, v in description.get("rewrites", {}).items()}
    out_dir.mkdir(parents=True, exist_ok=True)
    written = []
    for p in sources:
        text, records = expand_directives(texts[p.name], origin=p.stem)
        chosen.setdefault("directives", []).extend({"file": p.name, **r} for r in records)
        target = out_dir / p.name
        target.write_text(text)
        written.append(target)
    return written, chosen


def build(toml_path: Path, executable: Path, layer: Path, library: str = "ftc_hip", variant: str = "default",
          defines: tuple[str, ...] = (), keep: Path | None = None, source_root: Path | None = None) -> dict:
    """Generates the sources and builds `executable` against the host layer in `layer` (objects, modules
    and the compiler's path in `layer/compiler`) and lib<library>.so of this package."""
    fc = (layer / "compiler").read_text().strip()
    description = load_description(toml_path)
    defs = [*description.get("variants", {}).get(variant, {}).get("defines", ()), *defines]
    flags = ["-cpp", "-O3", "-fopenmp", "-ffp-contract=off", *(f"-D{d}" for d in defs)]
    with tempfile.TemporaryDirectory(prefix="ftc_app_") as tmp:
        work = keep or Path(tmp)
        sources, chosen = generate(toml_path, work / "src", source_root)
        objects = []
        run = lambda cmd: subprocess.run(cmd, cwd=work, check=True, capture_output=True, text=True)
        for src in [FORTRAN / "ftc_entries.f90", FORTRAN / "app" / "ftc_resident.f90", *sources]:
            obj = work / (src.stem + ".o")
            run([fc, *flags, "-I", str(layer), "-module-dir", str(work), "-c", str(src), "-o", str(obj)])
            objects.append(str(obj))
        executable.parent.mkdir(parents=True, exist_ok=True)
        # Found relative to the executable, so that a tree that moves keeps finding its library.
        rpath = "$ORIGIN/" + os.path.relpath(HERE, executable.parent)
        run([fc, "-fopenmp", *objects, str(layer / "scoped_memory.o"), str(layer / "resident_domain.o"),
             "-L", str(HERE), f"-l{library}", f"-Wl,-rpath,{rpath}", "-o", str(executable)])
    return chosen


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n", 1)[0])
    sub = ap.add_subparsers(dest="command", required=True)
    b = sub.add_parser("build", help="build the resident executable of an application description")
    b.add_argument("description", type=Path)
    b.add_argument("executable", type=Path)
    b.add_argument("--layer", type=Path, required=True, help="objects and modules of the Fortran host layer")
    b.add_argument("--library", default="ftc_hip")
    b.add_argument("--variant", default="default")
    b.add_argument("-D", dest="defines", action="append", default=[])
    g = sub.add_parser("generate", help="write the generated sources only")
    g.add_argument("description", type=Path)
    g.add_argument("out", type=Path)
    for p in (b, g):
        p.add_argument("--source-root", type=Path, default=None,
                       help="where the description's sources are (default: its own directory)")
    args = ap.parse_args(argv)
    root = args.source_root.resolve() if args.source_root else None
    if args.command == "build":
        chosen = build(args.description.resolve(), args.executable.resolve(), args.layer.resolve(), args.library,
                       args.variant, tuple(args.defines), source_root=root)
    else:
        _, chosen = generate(args.description.resolve(), args.out, root)
    for name, qualified in chosen["spliced"]:
        print(f"entry   {qualified} ({name})")
    for name, qualified in chosen["native"]:
        print(f"host    {qualified} ({name})")
    return 0


if __name__ == "__main__":
    sys.exit(main())
