package mopsgpu

// #include "mops.h"
import "C"

import (
	"errors"
	"runtime"

	"github.com/hrautila/matrix"
)

func imin(a, b int) int {
	if a < b {
		return a
	}
	return b
}

// DecomposeCHOL: A = L*L.T (LOWER) or U.T*U in place (chol.go:116-127). A
// NaN square root panics with the reference's message, "illegal value at
// <global index>: <value>" (chol.go:33-36).
func DecomposeCHOL(A *matrix.FloatMatrix, flags Flags, nb int) (*matrix.FloatMatrix, error) {
	if A.Cols() != A.Rows() {
		return nil, errors.New("A not a square matrix")
	}
	if A.Rows() == 0 {
		return A, nil
	}
	var pn runtime.Pinner
	defer pn.Unpin()
	a := view(&pn, A)
	rc := C.mops_dpotrf(&a, C.int(flags), C.int(A.Rows()), C.int(nb))
	if rc > 0 {
		panic(C.GoString(C.mops_last_error()))
	}
	mustOK(rc)
	return A, nil
}

// SolveCHOL: B = A.-1*B from the Cholesky factor (chol.go:147-158).
func SolveCHOL(B, A *matrix.FloatMatrix, flags Flags) {
	if A.Rows() == 0 || B.Cols() == 0 {
		return
	}
	var pn runtime.Pinner
	defer pn.Unpin()
	b, a := view(&pn, B), view(&pn, A)
	mustOK(C.mops_dpotrs(&b, &a, C.int(flags), C.int(A.Rows()), C.int(B.Cols())))
}

// DecomposeLU: A = P*L*U with partial pivoting in place (lu.go:257-273).
// pivots receive the 0-based global pivot rows (LAPACK ipiv - 1); ties go to
// the lowest row.
func DecomposeLU(A *matrix.FloatMatrix, pivots []int, nb int) (*matrix.FloatMatrix, error) {
	mlen := imin(A.Rows(), A.Cols())
	if len(pivots) < mlen {
		return A, errors.New("pivot array < min(A.Rows(),A.Cols())")
	}
	for k := range pivots {
		pivots[k] = 0
	}
	if mlen == 0 {
		return A, nil
	}
	var pn runtime.Pinner
	defer pn.Unpin()
	a := view(&pn, A)
	mustOK(C.mops_dgetrf(&a, pivotPtr(pivots), C.int(len(pivots)), C.int(A.Rows()),
		C.int(A.Cols()), C.int(nb)))
	return A, nil
}

// SolveLU: B = A.-1*B (A.-T*B with TRANSA) from DecomposeLU (lu.go:320-334).
// The TRANSA branch solves A.T*X = B (the reference applies the pivots
// first there, which does not).
func SolveLU(B, A *matrix.FloatMatrix, pivots []int, flags Flags) error {
	if A.Rows() == 0 || B.Cols() == 0 {
		return nil
	}
	var pn runtime.Pinner
	defer pn.Unpin()
	b, a := view(&pn, B), view(&pn, A)
	mustOK(C.mops_dgetrs(&b, &a, pivotPtr(pivots), C.int(len(pivots)), C.int(flags),
		C.int(A.Rows()), C.int(B.Cols())))
	return nil
}

// DecomposeLUnoPiv: A = L*U without pivoting in place (lu.go:291-299).
func DecomposeLUnoPiv(A *matrix.FloatMatrix, nb int) (*matrix.FloatMatrix, error) {
	if imin(A.Rows(), A.Cols()) == 0 {
		return A, nil
	}
	var pn runtime.Pinner
	defer pn.Unpin()
	a := view(&pn, A)
	mustOK(C.mops_dgetrf_nopiv(&a, C.int(A.Rows()), C.int(A.Cols()), C.int(nb)))
	return A, nil
}

// DecomposeLDLnoPiv: A = L*D*L.T (LOWER) or U*D*U.T without pivoting
// (ldlnp.go:211-241). W is required for blocked invocations, as in the
// reference, although the GPU keeps its own workspace.
func DecomposeLDLnoPiv(A, W *matrix.FloatMatrix, flags Flags, nb int) (*matrix.FloatMatrix, error) {
	if A.Cols() != A.Rows() {
		return nil, errors.New("A not a square matrix")
	}
	if !(A.Cols() < nb || nb == 0) && W == nil {
		return nil, errors.New("No workspace for blocking invocation")
	}
	if A.Rows() == 0 {
		return A, nil
	}
	var pn runtime.Pinner
	defer pn.Unpin()
	a := view(&pn, A)
	mustOK(C.mops_dsytrf_nopiv(&a, C.int(flags), C.int(A.Rows()), C.int(nb)))
	return A, nil
}

// SolveLDLnoPiv: B = A.-1*B from DecomposeLDLnoPiv (ldlnp.go:251-264).
func SolveLDLnoPiv(B, A *matrix.FloatMatrix, flags Flags) {
	if A.Rows() == 0 || B.Cols() == 0 {
		return
	}
	var pn runtime.Pinner
	defer pn.Unpin()
	b, a := view(&pn, B), view(&pn, A)
	mustOK(C.mops_dsytrs_nopiv(&b, &a, C.int(flags), C.int(A.Rows()), C.int(B.Cols())))
}

// DecomposeLDL: A = P.T*L*D*L.T*P (LOWER) or P.T*U*D*U.T*P with 1x1 diagonal
// pivots (ldl.go:439-460). W is not read (the GPU keeps its own workspace).
// A zero pivot gives the reference's error (or panics under
// SetPanicOnError).
func DecomposeLDL(A, W *matrix.FloatMatrix, ipiv []int, flags Flags, nb int) (*matrix.FloatMatrix, error) {
	if A.Cols() != A.Rows() {
		return nil, errors.New("A not a square matrix")
	}
	for k := range ipiv {
		ipiv[k] = 0
	}
	if A.Rows() == 0 {
		return A, nil
	}
	if len(ipiv) < A.Rows() {
		panic("runtime error: index out of range") // the reference's slice access
	}
	var pn runtime.Pinner
	defer pn.Unpin()
	a := view(&pn, A)
	rc := C.mops_dsytrf_ldl(&a, pivotPtr(ipiv), C.int(len(ipiv)), C.int(flags), C.int(A.Rows()),
		C.int(nb))
	if rc > 0 {
		return A, onError(C.GoString(C.mops_last_error()))
	}
	mustOK(rc)
	return A, nil
}

// SolveLDL: B = A.-1*B from DecomposeLDL (ldl.go:486-511).
func SolveLDL(B, A *matrix.FloatMatrix, ipiv []int, flags Flags) {
	if A.Rows() == 0 || B.Cols() == 0 {
		return
	}
	var pn runtime.Pinner
	defer pn.Unpin()
	b, a := view(&pn, B), view(&pn, A)
	mustOK(C.mops_dsytrs_ldl(&b, &a, pivotPtr(ipiv), C.int(len(ipiv)), C.int(flags),
		C.int(A.Rows()), C.int(B.Cols())))
}

// DecomposeBK: A = L*D*L.T (LOWER) or U*D*U.T (UPPER) with Bunch-Kaufman
// pivoting (ldlbk.go:1114-1140); ipiv in LAPACK's format. W is checked as
// the reference does and not read.
func DecomposeBK(A, W *matrix.FloatMatrix, ipiv []int, flags Flags, nb int) (*matrix.FloatMatrix, error) {
	for k := range ipiv {
		ipiv[k] = 0
	}
	need := nb + 1
	if A.Cols() < nb || nb == 0 {
		need = 2
	}
	if W == nil || W.Cols() < need || W.Rows() < A.Rows() {
		return nil, errors.New("Workspace too small")
	}
	if A.Rows() == 0 || flags&(LOWER|UPPER) == 0 {
		return A, nil
	}
	if len(ipiv) < A.Rows() {
		panic("runtime error: index out of range")
	}
	var pn runtime.Pinner
	defer pn.Unpin()
	a := view(&pn, A)
	mustOK(C.mops_dsytrf_bk(&a, pivotPtr(ipiv), C.int(len(ipiv)), C.int(flags), C.int(A.Rows()),
		C.int(nb)))
	return A, nil
}

// SolveBK: B = A.-1*B from DecomposeBK (ldlbks.go:303-316); the UPPER
// branch follows LAPACK's dsytrs.
func SolveBK(B, A *matrix.FloatMatrix, ipiv []int, flags Flags) {
	if A.Rows() == 0 || B.Cols() == 0 {
		return
	}
	var pn runtime.Pinner
	defer pn.Unpin()
	b, a := view(&pn, B), view(&pn, A)
	mustOK(C.mops_dsytrs_bk(&b, &a, pivotPtr(ipiv), C.int(len(ipiv)), C.int(flags),
		C.int(A.Rows()), C.int(B.Cols())))
}

// InverseTrm: the LOWER (else upper) triangle of A replaced by that of its
// inverse, UNIT for a unit diagonal (trinv.go:219-245).
func InverseTrm(A *matrix.FloatMatrix, flags Flags, nb int) (*matrix.FloatMatrix, error) {
	if A.Rows() == 0 {
		return A, nil
	}
	var pn runtime.Pinner
	defer pn.Unpin()
	a := view(&pn, A)
	mustOK(C.mops_dinverse_trm(&a, C.int(flags), C.int(A.Rows()), C.int(nb)))
	return A, nil
}

// Precisions of SolveIR.
const (
	PREC_FP64 = 0
	PREC_FP32 = 1
	PREC_BF16 = 2
)

// IRReport is the outcome of SolveIR (mops_ir_report).
type IRReport struct {
	Iterations    int
	Converged     bool
	BackwardError float64
	FirstError    float64
	FactorMs      float64
	RefineMs      float64
}

// SolveIR: X = A.-1*B by LU in precision prec with FP64 iterative
// refinement to a normwise backward error <= tol (no counterpart in
// matops). A and B are not modified.
func SolveIR(X, A, B *matrix.FloatMatrix, prec int, tol float64, maxit int) (IRReport, error) {
	var rep C.mops_ir_report
	n := A.Rows()
	if A.Cols() != n || B.Rows() != n || X.Rows() != n || X.Cols() != B.Cols() {
		return IRReport{}, errors.New("SolveIR: size mismatch")
	}
	var pn runtime.Pinner
	defer pn.Unpin()
	a, b, x := view(&pn, A), view(&pn, B), view(&pn, X)
	if err := status(C.mops_dgesv_ir(&a, &b, &x, C.int(n), C.int(B.Cols()), C.int(prec),
		C.double(tol), C.int(maxit), &rep)); err != nil {
		return IRReport{}, err
	}
	return IRReport{int(rep.iterations), rep.converged != 0, float64(rep.backward_error),
		float64(rep.first_error), float64(rep.factor_ms), float64(rep.refine_ms)}, nil
}

// CholeskyAdjoint: Lbar (d loss / dL, lower triangle) is overwritten with
// the symmetric d loss / dA for L = chol(A) (no counterpart in matops).
func CholeskyAdjoint(L, Lbar *matrix.FloatMatrix, nb int) error {
	n := L.Rows()
	if L.Cols() != n || Lbar.Rows() != n || Lbar.Cols() != n {
		return errors.New("L, Lbar size mismatch")
	}
	if n == 0 {
		return nil
	}
	var pn runtime.Pinner
	defer pn.Unpin()
	l, lb := view(&pn, L), view(&pn, Lbar)
	return status(C.mops_dpotrf_adjoint(&l, &lb, C.int(n), C.int(nb)))
}
