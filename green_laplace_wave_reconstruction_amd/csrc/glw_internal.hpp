// Internal helpers shared by the glw translation units.
#pragma once

#include <string>

// Record `msg` as the thread's last error (glw_last_error) and return `code`.
int glw_report_error(int code, const std::string& msg);
