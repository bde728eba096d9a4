package mopsgpu

// #include "mops.h"
import "C"

import (
	"errors"
	"math"
	"runtime"
	"sync"

	"github.com/hrautila/matrix"
)

// Work split of Mult / MultSym over NumWorkers goroutines, as the reference
// does it (mmult.go:105-188): block starts rounded down to multiples of 4,
// the worker grid chosen from the shape of C.
func blockIndex4(i, r, sz int) int {
	if i == r {
		return sz
	}
	v := i * sz / r
	return v - (v & 0x3)
}

func divideWork(rows, cols, workers int) (colWorkers, rowWorkers int) {
	nwsqrt := int(math.Sqrt(float64(workers)))
	switch {
	case workers == 2 || workers&0x1 != 0:
		if cols > rows {
			return workers, 1
		}
		return 1, workers
	case nwsqrt*nwsqrt == workers:
		return nwsqrt, nwsqrt
	case cols > rows:
		return workers / 2, 2
	default:
		return 2, workers / 2
	}
}

func scheduleWork(colworks, rowworks, cols, rows int, worker func(cs, ce, rs, re int)) {
	var wg sync.WaitGroup
	for k := 0; k < colworks; k++ {
		cs, ce := blockIndex4(k, colworks, cols), blockIndex4(k+1, colworks, cols)
		for l := 0; l < rowworks; l++ {
			rs, re := blockIndex4(l, rowworks, rows), blockIndex4(l+1, rowworks, rows)
			wg.Add(1)
			go func(cs, ce, rs, re int) {
				defer wg.Done()
				worker(cs, ce, rs, re)
			}(cs, ce, rs, re)
		}
	}
	wg.Wait()
}

// Mult: C = beta*C + alpha*op(A)*op(B) (mmult.go:198-256).
func Mult(Cm, A, B *matrix.FloatMatrix, alpha, beta float64, flags Flags) error {
	var ok, empty bool
	ar, ac := A.Size()
	br, bc := B.Size()
	cr, cc := Cm.Size()
	switch flags & (TRANSA | TRANSB) {
	case TRANSA | TRANSB:
		empty = ac == 0 || br == 0
		ok = cr == ac && cc == br && ar == bc
	case TRANSA:
		empty = ac == 0 || bc == 0
		ok = cr == ac && cc == bc && ar == br
	case TRANSB:
		empty = ar == 0 || br == 0
		ok = cr == ar && cc == br && ac == bc
	default:
		empty = ar == 0 || bc == 0
		ok = cr == ar && cc == bc && ac == br
	}
	if empty {
		return nil
	}
	if !ok {
		return errors.New("Mult: size mismatch")
	}
	psize := int64(Cm.NumElements()) * int64(A.Cols())
	P := A.Cols()
	if flags&TRANSA != 0 {
		P = A.Rows()
	}
	var pn runtime.Pinner
	defer pn.Unpin()
	c, a, b := view(&pn, Cm), view(&pn, A), view(&pn, B)
	worker := func(cs, ce, rs, re int) {
		cl, al, bl := c, a, b // per goroutine copies of the view structs
		mustOK(C.mops_dmult_mm_blocked3(&cl, &al, &bl, C.double(alpha), C.double(beta),
			C.int(flags), C.int(P), C.int(cs), C.int(ce), C.int(rs), C.int(re),
			C.int(vpLen), C.int(nB), C.int(mB)))
	}
	if nWorker <= 1 || psize <= limitOne {
		worker(0, cc, 0, cr)
		return nil
	}
	colworks, rowworks := divideWork(cr, cc, nWorker)
	scheduleWork(colworks, rowworks, cc, cr, worker)
	return nil
}

// MultSym: C = beta*C + alpha*A*B (LEFT) or alpha*B*A (RIGHT), A symmetric
// with one triangle stored (mmult.go:266-326).
func MultSym(Cm, A, B *matrix.FloatMatrix, alpha, beta float64, flags Flags) error {
	var ok, empty bool
	ar, ac := A.Size()
	br, bc := B.Size()
	cr, cc := Cm.Size()
	switch flags & (TRANSA | TRANSB) {
	case TRANSA | TRANSB:
		empty = ac == 0 || br == 0
		ok = ar == ac && cr == ac && cc == br && ar == bc
	case TRANSA:
		empty = ac == 0 || bc == 0
		ok = ar == ac && cr == ac && cc == bc && ar == br
	case TRANSB:
		empty = ar == 0 || br == 0
		ok = ar == ac && cr == ar && cc == br && ac == bc
	default:
		empty = ar == 0 || bc == 0
		ok = ar == ac && cr == ar && cc == bc && ac == br
	}
	if empty {
		return nil
	}
	if !ok {
		return errors.New("MultSym: size mismatch")
	}
	psize := int64(Cm.NumElements()) * int64(A.Cols())
	var pn runtime.Pinner
	defer pn.Unpin()
	c, a, b := view(&pn, Cm), view(&pn, A), view(&pn, B)
	worker := func(cs, ce, rs, re int) {
		cl, al, bl := c, a, b
		mustOK(C.mops_dmult_symm_blocked2(&cl, &al, &bl, C.double(alpha), C.double(beta),
			C.int(flags), C.int(ac), C.int(cs), C.int(ce), C.int(rs), C.int(re),
			C.int(vpLen), C.int(nB), C.int(mB)))
	}
	if nWorker <= 1 || psize <= limitOne {
		worker(0, cc, 0, cr)
		return nil
	}
	colworks, rowworks := divideWork(cr, cc, nWorker)
	scheduleWork(colworks, rowworks, cc, cr, worker)
	return nil
}

// MultTrm: B = alpha*op(A)*B (LEFT) or alpha*B*op(A) (RIGHT), A triangular
// (mmult.go:338-372).
func MultTrm(B, A *matrix.FloatMatrix, alpha float64, flags Flags) error {
	ok, empty := true, false
	br, bc := B.Size()
	ar, ac := A.Size()
	switch flags & (LEFT | RIGHT) {
	case LEFT:
		empty = br == 0 || bc == 0
		ok = br == ac && ac == ar
	case RIGHT:
		empty = bc == 0 || br == 0
		ok = bc == ar && ac == ar
	}
	if empty {
		return nil
	}
	if !ok {
		return onError("A, B size mismatch")
	}
	E := bc
	if flags&RIGHT != 0 {
		E = br
	}
	var pn runtime.Pinner
	defer pn.Unpin()
	b, a := view(&pn, B), view(&pn, A)
	mustOK(C.mops_dmmat_trmm_blk(&b, &a, C.double(alpha), C.int(flags), C.int(ac), 0, C.int(E),
		C.int(nB)))
	return nil
}

// SolveTrm: B = alpha*op(A)^-1*B (LEFT) or alpha*B*op(A)^-1 (RIGHT)
// (mmult.go:388-422).
func SolveTrm(B, A *matrix.FloatMatrix, alpha float64, flags Flags) error {
	ok, empty := true, false
	br, bc := B.Size()
	ar, ac := A.Size()
	switch flags & (LEFT | RIGHT) {
	case LEFT:
		empty = br == 0
		ok = br == ac && ac == ar
	case RIGHT:
		empty = bc == 0
		ok = bc == ar && ac == ar
	}
	if empty {
		return nil
	}
	if !ok {
		return onError("A, B size mismatch")
	}
	E := bc
	if flags&RIGHT != 0 {
		E = br
	}
	var pn runtime.Pinner
	defer pn.Unpin()
	b, a := view(&pn, B), view(&pn, A)
	mustOK(C.mops_dmmat_solve_blk(&b, &a, C.double(alpha), C.int(flags), C.int(ac), 0, C.int(E),
		C.int(nB)))
	return nil
}

// RankUpdateSym: C = beta*C + alpha*A*A.T (A.T*A with TRANSA) on one
// triangle (mmult.go:426-445).
func RankUpdateSym(Cm, A *matrix.FloatMatrix, alpha, beta float64, flags Flags) error {
	if Cm.Rows() != Cm.Cols() {
		return onError("C not a square matrix")
	}
	P := A.Cols()
	if flags&TRANSA != 0 {
		P = A.Rows()
	}
	var pn runtime.Pinner
	defer pn.Unpin()
	c, a := view(&pn, Cm), view(&pn, A)
	mustOK(C.mops_dmmat_rank_blk(&c, &a, C.double(alpha), C.double(beta), C.int(flags),
		C.int(P), 0, C.int(Cm.Rows()), C.int(vpLen), C.int(nB)))
	return nil
}

// RankUpdate2Sym: C = beta*C + alpha*A*B.T + alpha*B*A.T on one triangle
// (mmult.go:453-476).
func RankUpdate2Sym(Cm, A, B *matrix.FloatMatrix, alpha, beta float64, flags Flags) error {
	if Cm.Rows() != Cm.Cols() {
		return onError("C not a square matrix")
	}
	P := A.Cols()
	if flags&TRANSA != 0 {
		P = A.Rows()
	}
	var pn runtime.Pinner
	defer pn.Unpin()
	c, a, b := view(&pn, Cm), view(&pn, A), view(&pn, B)
	mustOK(C.mops_dmmat_rank2_blk(&c, &a, &b, C.double(alpha), C.double(beta), C.int(flags),
		C.int(P), 0, C.int(Cm.Rows()), C.int(vpLen), C.int(nB)))
	return nil
}

// UpdateTrm: C = beta*C + alpha*op(A)*op(B) on one triangle
// (mmult.go:501-520).
func UpdateTrm(Cm, A, B *matrix.FloatMatrix, alpha, beta float64, flags Flags) error {
	if Cm.Rows() != Cm.Cols() {
		return onError("C not a square matrix")
	}
	P := A.Cols()
	if flags&TRANSA != 0 {
		P = A.Rows()
	}
	var pn runtime.Pinner
	defer pn.Unpin()
	c, a, b := view(&pn, Cm), view(&pn, A), view(&pn, B)
	mustOK(C.mops_dmmat_trmupd_blk(&c, &a, &b, C.double(alpha), C.double(beta), C.int(flags),
		C.int(P), 0, C.int(Cm.Rows()), C.int(vpLen), C.int(nB)))
	return nil
}
