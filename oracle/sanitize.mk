# Host-side AddressSanitizer + UndefinedBehaviorSanitizer run of the exact C
# oracle (test infrastructure only; CPU, no GPU code involved):
#   make -f oracle/sanitize.mk
CC ?= gcc
HERE := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
BIN := $(HERE)_build/sanitize_check

all: run

$(BIN): $(HERE)c/tfhe_exact.c $(HERE)c/sanitize_check.c
	mkdir -p $(HERE)_build
	$(CC) -std=c11 -O1 -g -fno-omit-frame-pointer -fsanitize=address,undefined -fno-sanitize-recover=all -o $@ $^

run: $(BIN)
	ASAN_OPTIONS=detect_leaks=1:abort_on_error=0 UBSAN_OPTIONS=print_stacktrace=1 $(BIN)

.PHONY: all run
