"""Trained-model files: ARFF data and alpha files, serialized grids.

Host-side readers that let a model trained by the reference (a serialized
grid plus an alpha ARFF file) be admitted into the served loop or bound to
the operator:

* `read_arff` follows `ARFFTools::readARFF` / `readARFFSize`
  (datadriven/src/sgpp/datadriven/tools/ARFFTools.cpp:72-197): every line that
  contains '%' or '@' anywhere is skipped, and so is an empty line; the first
  data line fixes the column count (commas + 1) and every later line must
  match it; values are comma-separated and parsed like C `atof` (the longest
  numeric prefix after leading blanks, 0.0 when there is none); with targets,
  the last column is the target. `instance_cutoff`, `selected_cols` and
  `selected_targets` (|target - t| < 0.001) act as in the reference.
* `read_alpha` reads an alpha file: an ARFF file with one column and no
  targets (the reference's `LearnedKnowledgeFormatter` writes them with the
  attribute `alpha`).
* `write_alpha` writes one in the layout of the reference's
  `LearnedKnowledgeFormatter.__writeAlphaARFF`
  (datadriven/python/learner/formatter/LearnedKnowledgeFormatter.py:97-105):
  relation line, the attribute `alpha`, values as "%1.20f" (fixed point, so
  magnitudes below 1e-20 are written as zero); `save_model` writes the grid
  (format version 9) and its alpha file.
* `load_model` reads a serialized grid (`Grid::unserialize`, format versions
  1 to 9; `capi.Grid.unserialize`) and its alpha file, and checks that they
  agree in size.
"""
import re

import numpy as np

_ATOF = re.compile(r"\s*([+-]?(?:\d+\.?\d*(?:[eE][+-]?\d+)?|\.\d+(?:[eE][+-]?\d+)?|"
                   r"inf(?:inity)?|nan))", re.IGNORECASE)


def _atof(tok):
    m = _ATOF.match(tok)
    return float(m.group(1)) if m else 0.0


class ARFFError(ValueError):
    """Malformed ARFF content (the reference's file_exception)."""


def _text_of(source):
    """The source itself when it holds a newline, else the named file's text."""
    if "\n" in source:
        return source
    with open(source) as f:
        return f.read()


def _data_lines(text):
    for line in text.split("\n"):
        line = line.rstrip("\r")
        if "%" in line or "@" in line or line == "":
            continue
        yield line


def read_arff(source, has_targets=True, instance_cutoff=None, selected_cols=(),
              selected_targets=()):
    """Returns (data [n x dim] float64, targets [n] float64 or None).

    source: a path or the file's text (text when it contains a newline)."""
    lines = list(_data_lines(_text_of(source)))
    dim = 0
    n_inst = 0
    for k, line in enumerate(lines):
        commas = line.count(",")
        if dim == 0:
            dim = commas + (0 if has_targets else 1)
        elif dim - commas != (0 if has_targets else 1):
            raise ARFFError("readARFFSize: Columns missing in line %d" % n_inst)
        if has_targets and selected_targets:
            cl = _atof(line[line.rfind(",") + 1:])
            if any(abs(cl - t) < 0.001 for t in selected_targets):
                n_inst += 1
        else:
            n_inst += 1
    if selected_cols:
        if max(selected_cols) >= dim:
            raise ARFFError("readCSVPartial invalid col selection")
        out_dim = len(selected_cols)
    else:
        out_dim = dim
    if instance_cutoff is not None:
        n_inst = min(n_inst, instance_cutoff)
    data = np.zeros((n_inst, out_dim))
    targets = np.zeros(n_inst) if has_targets else None
    row = 0
    for line in lines:
        if row >= n_inst:
            break
        vals = [_atof(t) for t in line.split(",")]
        if has_targets:
            cl = vals.pop()
            if selected_targets and not any(abs(cl - t) < 0.001 for t in selected_targets):
                continue
            targets[row] = cl
        if selected_cols:
            data[row] = [vals[c] for c in selected_cols]
        else:
            data[row, :len(vals)] = vals[:out_dim]
        row += 1
    return data, targets


def read_alpha(source):
    """Coefficient vector of an alpha ARFF file (one column, no targets)."""
    data, _ = read_arff(source, has_targets=False)
    if data.ndim != 2 or data.shape[1] != 1:
        raise ARFFError("an alpha file has exactly one column")
    return data[:, 0].copy()


def load_model(grid_source, alpha_source):
    """(capi.Grid, alpha) of a serialized grid and its alpha file.

    grid_source / alpha_source: paths or the files' text."""
    from . import capi
    grid = capi.Grid.unserialize(_text_of(grid_source))
    alpha = read_alpha(alpha_source)
    if alpha.shape[0] != grid.size:
        raise ARFFError("alpha file holds %d coefficients, the grid %d points"
                        % (alpha.shape[0], grid.size))
    return grid, alpha


def write_alpha(path, alpha, relation=None):
    """Alpha ARFF file as the reference's formatter writes it."""
    alpha = np.asarray(alpha, dtype=np.float64).ravel()
    with open(path, "w") as f:
        f.write("@RELATION \"%s ALPHAFILE\"\n\n" % (relation if relation is not None else path))
        f.write("@ATTRIBUTE alpha NUMERIC\n")
        f.write("\n@DATA\n")
        for a in alpha:
            f.write("%1.20f\n" % a)


def save_model(grid_path, alpha_path, grid, alpha):
    """Serialized grid (format version 9) and its alpha file."""
    alpha = np.asarray(alpha, dtype=np.float64).ravel()
    if alpha.shape[0] != grid.size:
        raise ARFFError("alpha holds %d coefficients, the grid %d points"
                        % (alpha.shape[0], grid.size))
    with open(grid_path, "w") as f:
        f.write(grid.serialize())
    write_alpha(alpha_path, alpha)
