# The C restatement and its sanitizer harness (asan_check.c), built plain and
# with AddressSanitizer + UndefinedBehaviorSanitizer (host code only; gcc,
# no device code).  Test infrastructure, run by tests/test_oracle_sanitized.py:
#   make -f oracle/sanitize.mk
CC ?= gcc
HERE := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
SANFLAGS := -O1 -g -fno-omit-frame-pointer -fsanitize=address,undefined \
            -fno-sanitize-recover=all -march=x86-64-v3 -fopenmp -std=c11 -Wall

all: $(HERE)build/asan_check $(HERE)build/asan_check_plain

$(HERE)build/asan_check: $(HERE)restate.c $(HERE)asan_check.c
	mkdir -p $(HERE)build
	$(CC) $(SANFLAGS) -o $@ $^ -lm

$(HERE)build/asan_check_plain: $(HERE)restate.c $(HERE)asan_check.c
	mkdir -p $(HERE)build
	$(CC) -O3 -march=x86-64-v3 -fopenmp -std=c11 -Wall -o $@ $^ -lm

clean:
	rm -f $(HERE)build/asan_check $(HERE)build/asan_check_plain

.PHONY: all clean
