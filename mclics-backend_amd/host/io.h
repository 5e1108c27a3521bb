// Parameter file parser, GHRL coefficient setup and output writers of the
// host drop-in (reference src/io.cpp).
#pragma once
#include <cstdio>
#include <string>

#include "parameters.h"

namespace mc {

// read_input_file (src/io.cpp:81-253): whitespace-separated "key value"
// tokens, case-insensitive keys with the reference's aliases, '#' starts a
// comment that runs to the end of the line, anything after a value on its
// line is ignored, surface keys take an index first ("W 0 4").  Errors throw
// ConfigError with the reference's exit status: 5 unreadable file, 3 unknown
// key, 1 non-numeric value or nk != 1.  A null path returns the defaults.
Parameters read_input_file(const char *fname);

// setGHRL (src/io.cpp:277-302): lambda, mu, rho, nu, sigma from k11, k22,
// k33 and p0.
void setGHRL(Parameters &params);

void print_separator(const char *seg = "-", int count = 60);
void print_header(const char *title, int count = 60, const char *seg = "-");
void print_field(const char *label, const char *value);
void print_field(const char *label, double value);
void print_field(const char *label, int value);
void print_parameters(const Parameters &params);

// %g text of a float exactly as printf("%g", (double)v) writes it.
int format_g(char *buf, double v);

}  // namespace mc
