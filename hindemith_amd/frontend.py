"""Python 3 port of the reference's composition front end (PyAST.py, Templates.py).

The user's function is parsed with ``ast``; every node is visited children
first (PyAST.py:52-58) and mapped to semantic-model Filters by the same
dispatch rules as Templates.py:84-780, restricted to the Horn-Schunck path:

=====================  ===========================================  =====================
Python syntax          operand kinds                                Filter (Templates.py)
=====================  ===========================================  =====================
``a op b``             Array2D, Array2D                             Array2DOp (:103-120)
``a op s``/``s op a``  Array2D + Scalar                             Array2DScalarOp /
                                                                    ScalarArray2DOp (:141-196)
``a op b`` (1-D)       Array, Array / Scalar                        ArrayOp, ArrayScalarOp,
                                                                    ScalarArrayOp (:84-236)
``s op t``             Scalar, Scalar                               ScalarOp (:238-255)
``S * x``              Stencil, Array2D                             Stencil2DMVM (:601-617)
``A * x``              SpMat2D, Array2D                             SpmvDIA2D (:583-599)
``A * x``              csr_matrix, Array                            SpmvCSR (:619-636)
``A * x``              dia_matrix, Array                            SpmvDIA (:565-581)
``sum2d(a)``           Array2D                                      Array2DSum (:463-478)
``sum(a)``             Array                                        ArraySum (:446-461)
``sqrt(x)``            Scalar / Array / Array2D                     ScalarSqrt, ArraySqrt,
                                                                    Array2DSqrt (:480-529)
``x = expr``           per value kind                               Array2DCopy, ArrayCopy,
                                                                    ScalarCopy (:638-758)
``applyHSPreconditioner(D, diag0, diag1, offdiag, x0, x1)``          ApplyHSPreconditioner
                       Stencil, Array2D x 5 -> (Array2D, Array2D)   (:412-437)
``hs_matrix_powers(lam, Ix, Iy, r0, r1)``                           HSMatrixPowers
                       Scalar, Array2D x 4 -> (Array3D, Array3D)    (:277-300), STEPS planes
``V = <Array3D>``      Array3D                                      Array3DCopy (:698-730)
``warp_img2d(src, ox, oy)`` Array2D x 3 -> Array2D                  WarpImg2D (:257-274)
``set_brox_matrix(A, psi, alpha)`` statement, writes A's planes      SetBroxMatrix (:311-330)
``set_brox_redblack_matrix(A, psi, alpha)`` -> diagonal Array2D      SetBroxRedblackMatrix
``applyBroxPreconditioner(A, d0, d1, od, x0, x1)`` -> 2 Array2Ds     ApplyBroxPreconditioner
``red(x)`` / ``black(x)``  Array2D -> Array2D                         RedArray2D, BlackArray2D
``for i in range(N)``  N an int literal                             Iterator (:57-71)
``return a, b``                                                     Return (:760-780)
=====================  ===========================================  =====================

The semantic model and symbol table are serialised to the reference's XML wire
format (PyAST.py:61-133, without the xmllint reformatting) and handed to
libhmx (``hmx_compile``).  Behaviour kept from the reference: kwargs only,
literal loop bounds, ``defines`` must hold the tuning constants, results are
(u, v)-style tuples, and a returned kwarg array is updated in place (the
reference's returned arrays alias the caller's buffer, backend.cpp:186-276).
Reference defects fixed (each raised a KeyError, ran pdb or mis-ran before):

* SpmvCSR read ``shape`` from a Name node (Templates.py:628); the shape now
  comes from the symbol table, so ``A * x`` with a csr_matrix works.
* Numeric literals in arithmetic crashed (Templates.py:940-945 vs :1183); they
  now raise ``HindemithError`` with a pointer to the fix (pass np.float32).
* The record cache key was class names + ndarray shapes only
  (PyAST.py:158-173); it now also holds the function, kwarg names, dtypes,
  stencil/matrix sizes and scalar types.
* Unknown names dropped into ``pdb`` (Templates.py:1184); now an error.
"""
from __future__ import annotations

import ast
import itertools
import os
import textwrap
import threading

import numpy as np

from . import backend
from ._lib import HindemithError

__all__ = ["mycompile", "run_backend", "Record", "device_from_env", "set_plan_only"]

_name_counter = itertools.count()


def _unique_name():
    return f"_f{next(_name_counter)}"


# ---------------------------------------------------------------------------
# Semantic-model nodes (Templates.py:11-38): properties become XML attributes.

class Node:
    def __init__(self, name=None, **props):
        self.children = []
        self.properties = {}
        if name is not None:
            self.properties["name"] = name
        self.properties.update(props)

    @property
    def tag(self):
        return type(self).__name__


class SemanticModel(Node):
    pass


class PipeAndFilter(Node):
    pass


class Filter(Node):
    pass


class FunctionDef(Node):
    pass


class Iterator(Node):
    pass


class Name(Node):
    pass


class ScalarConstant(Node):
    pass


# Symbol kinds (Templates.py:920-1175)
class Scalar(Node):
    pass


class Stencil(Node):
    pass


class SpMat2D(Node):
    pass


class CSRMat(Node):
    pass


class SpMat(Node):
    pass


class Array(Node):
    pass


class Array2D(Node):
    pass


class Array3D(Node):
    pass


def _xml_escape(v):
    return (str(v).replace("&", "&amp;").replace("<", "&lt;").replace(">", "&gt;")
            .replace('"', "&quot;"))


def to_xml(node):
    """XMLVisitor.generic_visit (PyAST.py:75-88)."""
    s = "<" + node.tag
    for k, v in node.properties.items():
        s += f' {k}="{_xml_escape(v)}"'
    if node.children:
        s += ">" + "".join(to_xml(c) for c in node.children) + f"</{node.tag}>"
    else:
        s += "/>"
    return s


def symbol_table_xml(symtab):
    return "<SymbolTable>" + "".join(to_xml(v) for v in symtab.values()) + "</SymbolTable>"


def defines_xml(defines):
    out = "<ConstantTable>"
    for k, v in defines.items():
        out += to_xml(ScalarConstant(name=k, value=v))
    return out + "</ConstantTable>"


# ---------------------------------------------------------------------------
# Prototype: kwargs -> Header symbols (PyAST.py:11-41)

_OUT_OF_SCOPE_TYPES = {"complex64": "complex scalars"}


def _make_header(symtab, name, obj):
    cls = type(obj).__name__
    if cls == "SpMat2D":
        nd, d0, d1 = obj.data.shape
        symtab[name] = SpMat2D(name, shape=(d0 * d1, d0 * d1), source="Header", io="inout",
                               dtype="float32", ndiags=nd, dim0=d0, dim1=d1)
    elif cls == "Stencil":
        symtab[name] = Stencil(name, ndiags=len(obj.data), source="Header", io="inout",
                               dtype="float32")
    elif cls == "dia_matrix":  # PyAST.py:25-26
        n0, n1 = obj.shape
        if n0 != n1 or obj.data.shape[1] != n1:
            raise HindemithError(f"kwarg {name}: dia_matrix must be square with data of shape "
                                 "(ndiags, n) (SpMat.cpp:147-152 indexes it that way)")
        symtab[name] = SpMat(name, shape=tuple(obj.shape), source="Header", io="inout",
                             dtype=str(obj.dtype), ndiags=int(obj.data.shape[0]), dim=n0)
    elif cls in ("csr_matrix", "csr_array"):
        symtab[name] = CSRMat(name, shape=tuple(obj.shape), source="Header", io="inout",
                              dtype=str(obj.dtype), nnz=int(obj.nnz), dim=obj.shape[0],
                              dim0=obj.shape[0], dim1=obj.shape[1])
    elif cls == "ndarray":
        if obj.ndim == 1:
            symtab[name] = Array(name, source="Header", io="inout", dtype=str(obj.dtype),
                                 dim0=obj.shape[0], dim1=0, shape=tuple(obj.shape))
        elif obj.ndim == 2:
            symtab[name] = Array2D(name, source="Header", io="inout", dtype=str(obj.dtype),
                                   dim0=obj.shape[0], dim1=obj.shape[1],
                                   col_mjr=str(bool(obj.flags["F_CONTIGUOUS"]
                                                    and not obj.flags["C_CONTIGUOUS"])),
                                   shape=tuple(obj.shape))
        else:
            raise HindemithError(f"kwarg {name}: {obj.ndim}-D arrays (Array3D) are not on the "
                                 "Horn-Schunck path and are not supported by this backend")
    elif cls in ("float32", "float64", "int64"):
        symtab[name] = Scalar(name, shape=(), source="Header", io="inout", dtype=cls)
    elif cls in _OUT_OF_SCOPE_TYPES:
        raise HindemithError(f"kwarg {name}: {_OUT_OF_SCOPE_TYPES[cls]} is not supported by "
                             "this backend")
    # anything else is ignored, as in the reference; using it fails later


# ---------------------------------------------------------------------------
# Emit: bottom-up visit with the template rules

_BINOPS = {"Add", "Sub", "Mult", "Div"}
_OUT_OF_SCOPE_CALLS = {
    "lk_least_squares": "LKLeastSquares", "extract_snapshots": "ExtractSnapshots",
    "cgemm_batch": "CgemmBatch", "solve_batch": "SolveBatch", "gamma_weights": "GammaWeights",
    "inner_products": "InnerProducts",
}


class Emitter:
    def __init__(self, symtab, defines):
        self.symtab = symtab
        self.defines = defines
        self.vars = {}    # id(ast node) -> Name | [Name] | None
        self.nodes = {}   # id(ast node) -> [Filter/Iterator]
        self.model = SemanticModel()

    # -- helpers -----------------------------------------------------------
    def kind(self, v, where):
        if isinstance(v, list):
            raise HindemithError(f"{where}: a tuple cannot be used as an operand")
        if v is None:
            raise HindemithError(f"{where}: operand has no value")
        if "name" not in v.properties:
            raise HindemithError(
                f"{where}: numeric literals are not supported in @hm arithmetic (the reference "
                "crashes on them, Templates.py:940-945,1183); pass the constant as an "
                "np.float32 kwarg instead")
        nm = v.properties["name"]
        sym = self.symtab.get(nm)
        if sym is None:
            raise HindemithError(f"{where}: unknown symbol '{nm}' (not a supported kwarg and "
                                 "not assigned before use)")
        return sym.tag

    def create_copy(self, cls, name, like):
        """Templates.py CreateCopy: a new Inferred symbol with the properties of
        ``like`` unless ``name`` already exists."""
        cur = self.symtab.get(name)
        src = self.symtab[like.properties["name"]]
        if cur is None:
            var = cls(name)
            if cls is Scalar:
                var.properties.update(source="Inferred", dtype=src.properties["dtype"])
            else:
                var.properties = dict(src.properties)
                var.properties["name"] = name
                var.properties["source"] = "Inferred"
            self.symtab[name] = var
        else:
            want = cls.__name__
            if cur.tag != want:
                raise HindemithError(f"assignment to '{name}': it is a {cur.tag}, the value "
                                     f"is a {want}")
            if cls is not Scalar:
                for key in ("dim0", "dim1", "dim2", "col_mjr"):
                    if str(cur.properties.get(key)) != str(src.properties.get(key)):
                        raise HindemithError(f"assignment to '{name}': shape/layout differs "
                                             "from the value's")
        return Name(name)

    def visit(self, node):
        for child in ast.iter_child_nodes(node):
            self.visit(child)
        meth = getattr(self, "visit_" + type(node).__name__, None)
        if meth is not None:
            meth(node)

    def nodes_of(self, n, where):
        try:
            return self.nodes[id(n)]
        except KeyError:
            raise HindemithError(f"{where}: unsupported construct "
                                 f"'{ast.unparse(n) if hasattr(ast, 'unparse') else n}'")

    def var_of(self, n, where):
        if id(n) not in self.vars:
            raise HindemithError(f"{where}: unsupported expression "
                                 f"'{ast.unparse(n) if hasattr(ast, 'unparse') else n}'")
        return self.vars[id(n)]

    # -- leaves --------------------------------------------------------------
    def visit_Name(self, n):
        self.vars[id(n)] = Name(n.id)
        self.nodes[id(n)] = []

    def visit_Constant(self, n):  # Templates.py:73-82 (ast.Num) -> immediate
        self.vars[id(n)] = Name(val=str(n.value), dtype="float64")
        self.nodes[id(n)] = []

    def visit_Tuple(self, n):
        self.vars[id(n)] = [self.vars.get(id(e)) for e in n.elts]
        self.nodes[id(n)] = list(itertools.chain.from_iterable(self.nodes.get(id(e), [])
                                                               for e in n.elts))

    # -- expressions -----------------------------------------------------------
    def visit_BinOp(self, n):
        where = f"line {n.lineno}"
        lv = self.var_of(n.left, where)
        rv = self.var_of(n.right, where)
        lk, rk = self.kind(lv, where), self.kind(rv, where)
        op = type(n.op).__name__
        pair = (lk, rk)
        if pair in (("Stencil", "Array2D"), ("SpMat2D", "Array2D"), ("CSRMat", "Array"),
                    ("SpMat", "Array")):
            if op != "Mult":
                raise HindemithError(f"{where}: only '*' applies a {lk} to an array")
            fname = {"Stencil": "Stencil2DMVM", "SpMat2D": "SpmvDIA2D", "CSRMat": "SpmvCSR",
                     "SpMat": "SpmvDIA"}[lk]
            m = Filter(fname)
            if lk == "CSRMat":
                rows = int(self.symtab[lv.properties["name"]].properties["dim0"])
                nm = _unique_name()
                self.symtab[nm] = Array(nm, source="Inferred", io="inout",
                                        dtype=self.symtab[rv.properties["name"]].properties["dtype"],
                                        dim0=rows, dim1=0, shape=(rows,))
                newvar = Name(nm)
            elif lk == "SpMat":  # Templates.py:574: an Array shaped like x
                newvar = self.create_copy(Array, _unique_name(), rv)
            else:
                newvar = self.create_copy(Array2D, _unique_name(), rv)
        else:
            if op not in _BINOPS:
                raise HindemithError(f"{where}: operator {op} is not supported")
            table = {
                ("Array2D", "Array2D"): ("Array2DOp", Array2D, lv),
                ("Array2D", "Scalar"): ("Array2DScalarOp", Array2D, lv),
                ("Scalar", "Array2D"): ("ScalarArray2DOp", Array2D, rv),
                ("Array", "Array"): ("ArrayOp", Array, lv),
                ("Array", "Scalar"): ("ArrayScalarOp", Array, lv),
                ("Scalar", "Array"): ("ScalarArrayOp", Array, rv),
                ("Scalar", "Scalar"): ("ScalarOp", Scalar, lv),
            }
            if pair not in table:
                raise HindemithError(f"{where}: no operation for {lk} {op} {rk}")
            fname, cls, like = table[pair]
            if fname in ("Array2DOp", "ArrayOp"):
                a, b = (self.symtab[v.properties["name"]] for v in (lv, rv))
                for key in ("dim0", "dim1"):
                    if str(a.properties.get(key)) != str(b.properties.get(key)):
                        raise HindemithError(f"{where}: operands differ in shape")
            m = Filter(fname, op=op)
            newvar = self.create_copy(cls, _unique_name(), like)
        m.children += [newvar, lv, rv]
        self.vars[id(n)] = newvar
        self.nodes[id(n)] = self.nodes.get(id(n.left), []) + self.nodes.get(id(n.right), []) + [m]

    def visit_Call(self, n):
        fname = n.func.id if isinstance(n.func, ast.Name) else None
        where = f"line {n.lineno}"
        if fname in _OUT_OF_SCOPE_CALLS:
            raise HindemithError(f"{where}: {_OUT_OF_SCOPE_CALLS[fname]} ({fname}) is not on the "
                                 "Horn-Schunck path and is not supported by this backend")
        if fname == "applyHSPreconditioner":
            return self._apply_hs_preconditioner(n, where)
        if fname == "hs_matrix_powers":
            return self._hs_matrix_powers(n, where)
        if fname == "warp_img2d":
            return self._warp_img2d(n, where)
        if fname in ("set_brox_matrix", "set_brox_redblack_matrix"):
            return self._set_brox_matrix(n, where, fname == "set_brox_redblack_matrix")
        if fname == "applyBroxPreconditioner":
            return self._apply_brox_preconditioner(n, where)
        if fname in ("red", "black"):
            args = self._call_args(n, where, ["Array2D"])
            dst = self.create_copy(Array2D, _unique_name(), args[0])
            m = Filter("RedArray2D" if fname == "red" else "BlackArray2D")
            m.children += [dst] + args
            self.vars[id(n)] = dst
            self.nodes[id(n)] = self.nodes.get(id(n.args[0]), []) + [m]
            return
        rules = {("sum", "Array"): ("ArraySum", Scalar), ("sum2d", "Array2D"): ("Array2DSum", Scalar),
                 ("sqrt", "Scalar"): ("ScalarSqrt", Scalar), ("sqrt", "Array"): ("ArraySqrt", Array),
                 ("sqrt", "Array2D"): ("Array2DSqrt", Array2D)}
        if fname not in ("sum", "sum2d", "sqrt"):
            return  # e.g. range(N): consumed by the For rule; unknown calls fail when used
        if len(n.args) != 1:
            raise HindemithError(f"{where}: {fname} takes one argument")
        src = self.var_of(n.args[0], where)
        k = self.kind(src, where)
        if (fname, k) not in rules:
            raise HindemithError(f"{where}: {fname}() of a {k} is not supported")
        filt, cls = rules[(fname, k)]
        newvar = self.create_copy(cls, _unique_name(), src)
        m = Filter(filt)
        m.children += [newvar, src]
        self.vars[id(n)] = newvar
        self.nodes[id(n)] = self.nodes.get(id(n.args[0]), []) + [m]

    def _call_args(self, n, where, kinds):
        if len(n.args) != len(kinds) or n.keywords:
            raise HindemithError(f"{where}: {n.func.id} takes {len(kinds)} positional arguments")
        out = []
        for a, want in zip(n.args, kinds):
            v = self.var_of(a, where)
            k = self.kind(v, where)
            if k != want:
                raise HindemithError(f"{where}: {n.func.id}: expected a {want}, got a {k}")
            out.append(v)
        return out

    def _same_shape(self, vs, where):
        ref = self.symtab[vs[0].properties["name"]].properties
        for v in vs[1:]:
            p = self.symtab[v.properties["name"]].properties
            if any(str(p.get(k)) != str(ref.get(k)) for k in ("dim0", "dim1", "col_mjr")):
                raise HindemithError(f"{where}: operands differ in shape or layout")

    def _apply_hs_preconditioner(self, n, where):
        """Templates.py:412-437: (y0, y1) = applyHSPreconditioner(D, diag0, diag1,
        offdiag, x0, x1); the results are new Array2Ds shaped like x0 and x1."""
        args = self._call_args(n, where, ["Stencil"] + ["Array2D"] * 5)
        self._same_shape(args[1:], where)
        y0 = self.create_copy(Array2D, _unique_name(), args[4])
        y1 = self.create_copy(Array2D, _unique_name(), args[5])
        m = Filter("ApplyHSPreconditioner")
        m.children += [y0, y1] + args
        self.vars[id(n)] = [y0, y1]
        self.nodes[id(n)] = list(itertools.chain.from_iterable(
            self.nodes.get(id(a), []) for a in n.args)) + [m]

    def _set_brox_matrix(self, n, where, redblack):
        """Templates.py:311-350: set_brox_matrix(A, psi, alpha) as a statement (it
        writes A's planes), or diag = set_brox_redblack_matrix(A, psi, alpha)."""
        args = self._call_args(n, where, ["SpMat2D", "Array2D", "Scalar"])
        A = self.symtab[args[0].properties["name"]].properties
        psi = self.symtab[args[1].properties["name"]].properties
        if (str(A["dim0"]), str(A["dim1"])) != (str(psi["dim0"]), str(psi["dim1"])):
            raise HindemithError(f"{where}: {n.func.id}: matrix planes and psi differ in shape")
        m = Filter("SetBroxRedblackMatrix" if redblack else "SetBroxMatrix")
        if redblack:
            diag = self.create_copy(Array2D, _unique_name(), args[1])
            m.children.append(diag)
            self.vars[id(n)] = diag
        else:
            self.vars[id(n)] = None
        m.children += args
        self.nodes[id(n)] = list(itertools.chain.from_iterable(
            self.nodes.get(id(a), []) for a in n.args)) + [m]

    def _apply_brox_preconditioner(self, n, where):
        """Templates.py:375-400: (y0, y1) = applyBroxPreconditioner(A, diag0, diag1,
        offdiag, x0, x1) with A an SpMat2D whose plane 2 is the diagonal."""
        args = self._call_args(n, where, ["SpMat2D"] + ["Array2D"] * 5)
        self._same_shape(args[1:], where)
        y0 = self.create_copy(Array2D, _unique_name(), args[4])
        y1 = self.create_copy(Array2D, _unique_name(), args[5])
        m = Filter("ApplyBroxPreconditioner")
        m.children += [y0, y1] + args
        self.vars[id(n)] = [y0, y1]
        self.nodes[id(n)] = list(itertools.chain.from_iterable(
            self.nodes.get(id(a), []) for a in n.args)) + [m]

    def _warp_img2d(self, n, where):
        """Templates.py:257-274: dst = warp_img2d(src, offx, offy), a new Array2D
        shaped like src."""
        args = self._call_args(n, where, ["Array2D"] * 3)
        self._same_shape(args, where)
        dst = self.create_copy(Array2D, _unique_name(), args[0])
        m = Filter("WarpImg2D")
        m.children += [dst] + args
        self.vars[id(n)] = dst
        self.nodes[id(n)] = list(itertools.chain.from_iterable(
            self.nodes.get(id(a), []) for a in n.args)) + [m]

    def _hs_matrix_powers(self, n, where):
        """Templates.py:277-300: (V0, V1) = hs_matrix_powers(lam, Ix, Iy, r0, r1),
        two Array3Ds of defines['STEPS'] planes shaped like r0."""
        args = self._call_args(n, where, ["Scalar"] + ["Array2D"] * 4)
        self._same_shape(args[1:], where)
        if "STEPS" not in self.defines:
            raise HindemithError(f"{where}: hs_matrix_powers needs defines['STEPS'] "
                                 "(HSMatrixPowers.cpp:153)")
        steps = int(self.defines["STEPS"])
        if steps < 1:
            raise HindemithError(f"{where}: defines['STEPS'] must be >= 1")
        r = self.symtab[args[3].properties["name"]].properties
        if r.get("col_mjr") == "True":
            raise HindemithError(f"{where}: hs_matrix_powers needs C-contiguous arrays")
        outs = []
        for _ in range(2):
            nm = _unique_name()
            self.symtab[nm] = Array3D(nm, source="Inferred", io="inout", dtype=r["dtype"],
                                      dim0=steps, dim1=r["dim0"], dim2=r["dim1"],
                                      col_mjr="False", shape=(steps, r["dim0"], r["dim1"]))
            outs.append(Name(nm))
        m = Filter("HSMatrixPowers")
        m.children += outs + args
        self.vars[id(n)] = outs
        self.nodes[id(n)] = list(itertools.chain.from_iterable(
            self.nodes.get(id(a), []) for a in n.args)) + [m]

    # -- statements ---------------------------------------------------------------
    def visit_Assign(self, n):
        where = f"line {n.lineno}"
        value = self.var_of(n.value, where)
        if len(n.targets) != 1:
            raise HindemithError(f"{where}: chained assignment is not supported")
        tgt = n.targets[0]
        if isinstance(value, list):
            if not isinstance(tgt, ast.Tuple) or len(tgt.elts) != len(value):
                raise HindemithError(f"{where}: tuple assignment needs a matching tuple target")
            targets = tgt.elts
        else:
            value = [value]
            targets = [tgt]
        out = list(self.nodes.get(id(n.value), []))
        for t, v in zip(targets, value):
            if not isinstance(t, ast.Name):
                raise HindemithError(f"{where}: only plain names can be assigned")
            k = self.kind(v, where)
            filt, cls = {"Array2D": ("Array2DCopy", Array2D), "Array": ("ArrayCopy", Array),
                         "Scalar": ("ScalarCopy", Scalar),
                         "Array3D": ("Array3DCopy", Array3D)}.get(k, (None, None))
            if filt is None:
                raise HindemithError(f"{where}: cannot assign a {k}")
            target = self.create_copy(cls, t.id, v)
            m = Filter(filt)
            m.children += [target, v]
            out.append(m)
            self.vars[id(n)] = target
        self.nodes[id(n)] = out

    def visit_For(self, n):
        where = f"line {n.lineno}"
        it = n.iter
        ok = (isinstance(it, ast.Call) and isinstance(it.func, ast.Name) and it.func.id == "range"
              and len(it.args) == 1 and isinstance(it.args[0], ast.Constant)
              and isinstance(it.args[0].value, int))
        if not ok:
            raise HindemithError(f"{where}: loops must be 'for i in range(N)' with N an integer "
                                 "literal (Templates.py:63)")
        if n.orelse:
            raise HindemithError(f"{where}: for/else is not supported")
        f = Iterator(ub=it.args[0].value)
        body = PipeAndFilter()
        for s in n.body:
            body.children += self.nodes_of(s, where)
        f.children.append(body)
        self.nodes[id(n)] = [f]

    def visit_Return(self, n):
        where = f"line {n.lineno}"
        if n.value is None:
            raise HindemithError(f"{where}: return needs a value")
        value = self.var_of(n.value, where)
        m = Filter("Return")
        for v in (value if isinstance(value, list) else [value]):
            self.kind(v, where)
            m.children.append(v)
        self.nodes[id(n)] = self.nodes.get(id(n.value), []) + [m]

    def visit_Expr(self, n):
        if isinstance(n.value, ast.Constant) and isinstance(n.value.value, str):
            self.nodes[id(n)] = []  # docstring
        elif (isinstance(n.value, ast.Call) and isinstance(n.value.func, ast.Name)
              and n.value.func.id == "set_brox_matrix"):
            self.nodes[id(n)] = self.nodes.get(id(n.value), [])  # Templates.py:311-330

    def visit_Pass(self, n):
        self.nodes[id(n)] = []

    def visit_FunctionDef(self, n):
        f = FunctionDef(n.name)
        body = PipeAndFilter()
        for s in n.body:
            body.children += self.nodes_of(s, f"line {getattr(s, 'lineno', '?')}")
        f.children.append(body)
        self.model.children.append(f)
        self.nodes[id(n)] = [f]


# ---------------------------------------------------------------------------
# Records (PyAST.py:135-196 cache + backend.cpp static maps)

_HM_ENV = ("HM_CPU0", "HM_CPU0_SUB1", "HM_GPU0", "HM_GPU01", "HM_GPU1")


def device_from_env(env=None):
    """Device selection of clhelp.cpp:43-93, as an index or an exception."""
    env = os.environ if env is None else env
    if env.get("HM_CPU0") or env.get("HM_CPU0_SUB1"):
        raise HindemithError("HM_CPU0 / HM_CPU0_SUB1 select an OpenCL CPU device; this backend "
                             "runs on MI355X GPUs only (set HM_GPU0)")
    if env.get("HM_GPU0"):
        return 0
    if env.get("HM_GPU01"):
        raise HindemithError("HM_GPU01 (two-device split) is not supported: this backend runs "
                             "one GPU per process (set HM_GPU0 or HM_GPU1)")
    if env.get("HM_GPU1"):
        return 1
    raise HindemithError("Specify target: set HM_GPU0 or HM_GPU1 (clhelp.cpp:87-90)")


_records = {}
_rec_ids = itertools.count()
_init_lock = threading.Lock()
_initialised = [None]


_plan_only = [False]


def set_plan_only(enable=True):
    """Compile records without a GPU (plan inspection only; runs fail)."""
    with _init_lock:
        _plan_only[0] = bool(enable)
        _initialised[0] = None
        _records.clear()


def _ensure_init():
    with _init_lock:
        dev = -1 if _plan_only[0] else device_from_env()
        if _initialised[0] != dev:
            backend.init(dev)
            _initialised[0] = dev


def _kw_signature(kwargs):
    sig = []
    for k, v in kwargs.items():
        cls = type(v).__name__
        ent = [k, cls]
        if isinstance(v, np.ndarray):
            ent += [v.shape, str(v.dtype), bool(v.flags["F_CONTIGUOUS"] and not v.flags["C_CONTIGUOUS"])]
        elif cls in ("Stencil",):
            ent += [len(v.data)]
        elif cls == "SpMat2D":
            ent += [v.data.shape]
        elif cls == "dia_matrix":
            ent += [v.shape, v.data.shape, str(v.dtype)]
        elif cls in ("csr_matrix", "csr_array"):
            ent += [v.shape, int(v.nnz), str(v.dtype)]
        sig.append(tuple(ent))
    return tuple(sig)


class Record:
    """A compiled program for one kwargs signature."""

    def __init__(self, rec_id, symtab, model_xml, symtab_xml, consts_xml):
        self.id = rec_id
        self.symtab = symtab
        self.model_xml = model_xml
        self.symtab_xml = symtab_xml
        self.consts_xml = consts_xml

    def plan(self):
        return backend.plan(self.id)

    def program_source(self, compile=False):
        return backend.program_source(self.id, compile)


def mycompile(source, kwargs, defines, fn_key=None):
    """PyAST.mycompile (PyAST.py:149-196): compile once per kwargs signature."""
    key = (fn_key if fn_key is not None else source, _kw_signature(kwargs))
    rec = _records.get(key)
    if rec is not None:
        return rec
    _ensure_init()
    symtab = {}
    for k, v in kwargs.items():
        _make_header(symtab, k, v)
    tree = ast.parse(textwrap.dedent(source))
    em = Emitter(symtab, defines)
    em.visit(tree)
    if not em.model.children:
        raise HindemithError("no function definition found in the decorated source")
    model_xml = to_xml(em.model)
    st_xml = symbol_table_xml(symtab)
    c_xml = defines_xml(defines)
    rid = next(_rec_ids)
    backend.xmlcompile(rid, st_xml, model_xml, c_xml)
    rec = Record(rid, symtab, model_xml, st_xml, c_xml)
    _records[key] = rec
    return rec


def run_backend(rec, kwargs):
    """PyAST.run_backend (PyAST.py:198-200): backend.run, keep the return value."""
    retval, _state = backend.run(rec.id, kwargs)
    return retval
