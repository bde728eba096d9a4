// Package mopsgpu is matops' dense FP64 path on an AMD MI355X GPU: the
// functions of github.com/hrautila/matops that this path covers, with the
// same signatures, view semantics and errors, computed by libmops.so (the
// C ABI in include/mops.h) instead of the calgo CPU kernels. A caller
// switches by importing this package in place of matops for these calls.
//
// Build: libmops.so is built by `python -m matops_amd.build` in this
// repository; the cgo flags below find it relative to this directory (set
// CGO_CFLAGS / CGO_LDFLAGS to override).
package mopsgpu

// #cgo CFLAGS: -I${SRCDIR}/../../include
// #cgo LDFLAGS: -L${SRCDIR}/../../matops_amd -lmops -Wl,-rpath,${SRCDIR}/../../matops_amd
// #include <stdint.h>
// #include "mops.h"
import "C"

import (
	"errors"
	"runtime"
	"unsafe"

	"github.com/hrautila/matrix"
)

// Flags keep the bit values of matops / calgo (calgo/cmops.go:19-31).
type Flags int

const (
	TRANSA  Flags = 0x1
	TRANSB  Flags = 0x2
	LOWER   Flags = 0x4
	UPPER   Flags = 0x8
	LEFT    Flags = 0x10
	RIGHT   Flags = 0x20
	UNIT    Flags = 0x40
	TRANS   Flags = 0x80
	NOTRANS Flags = 0
	NONE    Flags = 0
)

// Tuning state of mmult.go:33-44. The block sizes are CPU hints that the
// GPU accepts and ignores; they are still passed through.
var (
	vpLen        = 196
	nB           = 68
	mB           = 68
	nWorker      = 1
	limitOne     = int64(200 * 200 * 200)
	panicOnError = false
)

// SetPanicOnError (mmult.go:55-57).
func SetPanicOnError(newval bool) { panicOnError = newval }

// BlockingParams (mmult.go:63-67).
func BlockingParams(mb, nb, kb int) {
	vpLen = kb
	nB = nb
	mB = mb
}

// NumWorkers (mmult.go:69-73): Mult and MultSym split C over this many
// goroutines, each calling the kernel-level entry point on its block; the
// library keeps a stream and workspace per host thread, so the calls run
// concurrently.
func NumWorkers(newWorkers int) int {
	old := nWorker
	nWorker = newWorkers
	return old
}

func onError(msg string) error { // mmult.go:46-51
	if panicOnError {
		panic(msg)
	}
	return errors.New(msg)
}

// view returns the calgo mdata_t of a FloatMatrix view: pointer at element
// (0,0) and the parent's leading index (calgo/cmops.go:76-82). The struct
// is Go memory that holds a Go pointer when passed to C, which the cgo rules
// allow only for pinned memory: the caller's Pinner pins the view's first
// element (Unpin after the call).
func view(pn *runtime.Pinner, A *matrix.FloatMatrix) C.mops_mdata_t {
	var m C.mops_mdata_t
	ar := A.FloatArray()
	if len(ar) > 0 {
		pn.Pin(&ar[0])
		m.md = (*C.double)(unsafe.Pointer(&ar[0]))
	}
	m.step = C.int(A.LeadingIndex())
	return m
}

// status turns a return code into nil or the library's error text. The
// reference's C kernels cannot fail; on the GPU a runtime failure (a HIP
// error, a timed-out internal wait) is reported, and panics where the
// reference signature has no error to return.
func status(rc C.int) error {
	if rc == 0 {
		return nil
	}
	return errors.New(C.GoString(C.mops_last_error()))
}

func mustOK(rc C.int) {
	if err := status(rc); err != nil {
		panic("mopsgpu: " + err.Error())
	}
}

func pivotPtr(p []int) *C.int64_t {
	if len(p) == 0 {
		return nil
	}
	// Go int is 64 bits on the platforms the library runs on
	return (*C.int64_t)(unsafe.Pointer(&p[0]))
}
