#pragma once
// Version of the API this drop-in implements (core/version.h:7-9).
#define TIMEEVOLVER_VERSION 201
