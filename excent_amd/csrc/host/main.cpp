// main.cpp -- process entry for the bin/ programs; EXCENT_ENTRY names the library
// function (excent_main, excent_rmat_main, ...) this binary runs.
#include "excent.h"

int main(int argc, char** argv) { return EXCENT_ENTRY(argc, (const char* const*)argv); }
