// CPU oracle: relic/bin index restatement (filled in below).
#pragma once
