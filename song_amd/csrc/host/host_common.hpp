// Shared error reporting of the host library (song_host.h).
#pragma once

int host_fail(int code, const char *fmt, ...);
