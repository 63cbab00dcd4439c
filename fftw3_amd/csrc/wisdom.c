/*
 * Wisdom store placeholder: decisions are not persisted yet.
 */
#include <stdlib.h>
#include <string.h>
#include "engine.h"

char *fk_wisdom_export(void)
{
    const char *s = "(fftw3-amd-gfx950-wisdom)\n";
    char *r = (char *)malloc(strlen(s) + 1);
    strcpy(r, s);
    return r;
}

int fk_wisdom_import(const char *s)
{
    return s && strncmp(s, "(fftw3-amd-gfx950-wisdom", 24) == 0;
}

void fk_forget_wisdom(void) {}
