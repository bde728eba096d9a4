module github.com/hrautila/matops/mopsgpu

go 1.21

require github.com/hrautila/matrix v0.0.0
